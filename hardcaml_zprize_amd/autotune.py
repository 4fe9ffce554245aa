"""Size-class autotuner for the MSM window width (and the NTT tile width).

For each size class log2 N it times `window` candidates around the library's
default policy on the GPU and keeps the fastest; the result is a small JSON
table that `MsmWorkspace.run` jobs (per-job `window`) and the serving loop
consult.  `ZP_NTT_TILE` (4 or 8) selects the 4096-point NTT kernels' tile.

    python -m hardcaml_zprize_amd.autotune --min-log 12 --max-log 22 --out autotune.json
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
    Returns {log_n: (best_window, {window: seconds})}."""
    from .msm import MsmWorkspace

    ws = MsmWorkspace()
    out = {}
    try:
        for lg in log_sizes:
            base = default_window(lg)
            cands = candidates or sorted({min(20, max(8, base + d)) for d in (-2, -1, 0, 1, 2)})
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
    ap.add_argument("--min-log", type=int, default=12)
    ap.add_argument("--max-log", type=int, default=22)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_PATH)
    args = ap.parse_args()
    from . import msm as zmsm

    rng = np.random.default_rng(7)
    n = 1 << args.max_log
    ks = rng.integers(0, 2**64, size=(n, 4), dtype=np.uint64)
    ks[:, 3] &= np.uint64((1 << 60) - 1)
    dk = zmsm.DeviceBuffer(32 * n).upload(ks)
    dp = zmsm.DeviceBuffer(104 * n)
    zmsm.generate_points_device(dk.ptr, dp.ptr, n)
    ctx = zmsm.MsmContext((dp.ptr, n))
    dp.free()
    sc = rng.integers(0, 2**64, size=(n, 4), dtype=np.uint64)
    sc[:, 3] &= np.uint64((1 << 60) - 1)
    ds = zmsm.DeviceBuffer(32 * n).upload(sc)
    res = tune_msm(ctx, ds.ptr, range(args.min_log, args.max_log + 1), reps=args.reps)
    table = {str(k): v[0] for k, v in res.items()}
    detail = {str(k): {str(w): t for w, t in v[1].items()} for k, v in res.items()}
    with open(args.out, "w") as f:
        json.dump({"msm_window": table, "seconds": detail}, f, indent=1)
    print(json.dumps({"msm_window": table}))
    ctx.close()
    ds.free()
    dk.free()


if __name__ == "__main__":
    main()
