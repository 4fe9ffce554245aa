"""Size-class autotuner for the MSM window width (and the NTT tile width).

For each size class log2 N it times `window` candidates around the library's
default policy on the GPU and keeps the fastest; the result is a small JSON
table that `MsmWorkspace.run` jobs (per-job `window`) and the serving loop
consult.  `ZP_NTT_TILE` (4 or 8) selects the 64 x 64 NTT kernels' tile (set, it
also puts the 2^24 transform's second pass on the 64 x 64 row kernel).

    python -m hardcaml_zprize_amd.autotune --min-log 10 --max-log 22 --options precompute=auto
"""
import argparse
import json
import os
import time

import numpy as np

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune.json")


def default_window(log_n):
    """The library's built-in policy (msm.hip auto_window)."""
    return min(20, max(8, log_n - 6))


def load(path=DEFAULT_PATH):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): int(v) for k, v in json.load(f).get("msm_window", {}).items()}


def window_for(log_n, table=None):
    """Tuned window for this size class, 0 (= library default) if untuned."""
    table = load() if table is None else table
    return table.get(int(log_n), 0)


def tune_msm(ctx, d_scalars, log_sizes, candidates=None, reps=3):
    """ctx: MsmContext with >= 2^max(log_sizes) points; d_scalars: device scalars.
    Candidates per size class: 0 (the library's own choice, which includes
    the window-shifted table when ctx has one) and the default width +- 2.
    Returns {log_n: (best_window, {window: seconds})}."""
    from .msm import MsmWorkspace

    ws = MsmWorkspace()
    out = {}
    try:
        for lg in log_sizes:
            base = default_window(lg)
            cands = candidates or [0] + sorted({min(20, max(8, base + d)) for d in (-2, -1, 0, 1, 2)})
            times = {}
            for w in cands:
                ws.run([(ctx, 1 << lg, d_scalars, w)])  # warm-up (workspace growth)
                t0 = time.perf_counter()
                for _ in range(reps):
                    ws.run([(ctx, 1 << lg, d_scalars, w)])
                times[w] = (time.perf_counter() - t0) / reps
            best = min(times, key=times.get)
            out[lg] = (best, times)
    finally:
        ws.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--min-log", type=int, default=10)
    ap.add_argument("--max-log", type=int, default=22)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--options", default="precompute=auto",
                    help="msm_init options of the table tuned on (the batched form's shared table)")
    ap.add_argument("--out", default=DEFAULT_PATH)
    args = ap.parse_args()
    from . import msm as zmsm
    from . import workload

    rng = np.random.default_rng(7)
    n = 1 << args.max_log
    ctx, _, _, _ = workload.generate_table(rng, args.max_log, args.options, keep_ks=False)
    ds = zmsm.DeviceBuffer(32 * n).upload(workload.rand_scalars(rng, n))
    res = tune_msm(ctx, ds.ptr, range(args.min_log, args.max_log + 1), reps=args.reps)
    table = {str(k): v[0] for k, v in res.items()}
    detail = {str(k): {str(w): t for w, t in v[1].items()} for k, v in res.items()}
    with open(args.out, "w") as f:
        json.dump({"msm_window": table, "seconds": detail, "table_options": args.options,
                   "table_points": n, "note": "window 0 = the library's own choice (size-class default or the "
                                              "window-shifted table)"}, f, indent=1)
    print(json.dumps({"msm_window": table}), flush=True)
    ctx.close()
    ds.free()


if __name__ == "__main__":
    main()
