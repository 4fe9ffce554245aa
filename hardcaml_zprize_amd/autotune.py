"""Size-class autotuner: MSM window widths and NTT tiles.

MSM: for each size class log2 N it times `window` candidates around the
library's default policy on the GPU and keeps the fastest.  NTT: for each
size class of the two-pass transform (2^10 .. 2^23 points) it chooses the
column length (4096: the register-resident column kernel) and the sequences
per workgroup of both LDS passes (zp_ntt_set_tiles), once for single
transforms and once for the groups of a batch (B jobs of one size on
grid.y), by coordinate descent: the split among the built-in one +- 1, then
pass 1's tile, then pass 2's.  The result is a
small JSON table (`autotune.json`) that `MixedBatch`, the serving loop and
the bench apply (`apply_ntt_tiles`, per-job MSM `window`).  `ZP_NTT_TILE`
(4 or 8) selects the 64 x 64 register kernels' tile for 2^21 and above (set,
it also puts the 2^24 transform's second pass on the 64 x 64 row kernel).

    python -m hardcaml_zprize_amd.autotune --min-log 10 --max-log 22 --options precompute=auto
    python -m hardcaml_zprize_amd.autotune --ntt-only     # NTT tiles, MSM table kept
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


def _read(path):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def load(path=DEFAULT_PATH):
    """size class -> tuned MSM window (empty: no table)."""
    return {int(k): int(v) for k, v in _read(path).get("msm_window", {}).items()}


def load_ntt_tiles(path=DEFAULT_PATH):
    """{"single" | "batched": {log_n: (log_n1, log_t1, log_t2)}} (empty: no table)."""
    t = _read(path).get("ntt_tiles", {})
    return {mode: {int(k): tuple(int(x) for x in v) for k, v in t.get(mode, {}).items()}
            for mode in ("single", "batched") if mode in t}


def apply_ntt_tiles(tiles=None, path=DEFAULT_PATH):
    """Set the library's per-size-class NTT tiles from the table; returns the
    number of (mode, size class) entries applied."""
    from . import ntt as zntt

    tiles = load_ntt_tiles(path) if tiles is None else tiles
    count = 0
    for mode, per in tiles.items():
        for lg, (n1, t1, t2) in per.items():
            zntt.set_tiles(lg, mode == "batched", n1, t1, t2)
            count += 1
    return count


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


def _time_ntt(log_n, batched, choice, bufs, reps):
    """Seconds per transform (single) or per batch launch (batched) under the
    tile choice; bufs: device buffers of 2^log_n elements (one per batch job)."""
    from . import ntt as zntt
    from ._lib import call

    zntt.set_tiles(log_n, batched, *choice)
    if batched:
        b = zntt.NttBatch([x.ptr for x in bufs], [log_n] * len(bufs), [0] * len(bufs))
        try:
            b.launch(True)  # capture (with the tiles just set) + first replay
            call("zp_device_synchronize")
            t0 = time.perf_counter()
            for _ in range(reps):
                b.launch(True)
            call("zp_device_synchronize")
            return (time.perf_counter() - t0) / reps
        finally:
            b.close()
    plan = zntt.NttGpuDriver(log_n, num_buffers=1)
    try:
        plan.inplace_device(bufs[0].ptr)
        call("zp_device_synchronize")
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.inplace_device(bufs[0].ptr)
        call("zp_device_synchronize")
        return (time.perf_counter() - t0) / reps
    finally:
        plan.close()


def tune_ntt(log_sizes, batch_jobs=12, rng=None):
    """Coordinate descent per size class and mode (single, batched): the split
    among the balanced one +- 1 with the built-in tiles, then pass 1's tile
    over every valid width, then pass 2's.  Leaves the library on the chosen
    tiles.  Returns {mode: {log_n: (choice, {choice_str: seconds})}}."""
    from . import msm as zmsm
    from . import ntt as zntt
    from . import workload

    rng = rng or np.random.default_rng(11)
    out = {"single": {}, "batched": {}}
    for lg in log_sizes:
        bufs = [zmsm.DeviceBuffer(8 << lg).upload(workload.rand_ntt_input(rng, 1 << lg)) for _ in range(batch_jobs)]
        # enough launches for ~20 ms of work at ~8 B/element/us
        reps = max(5, min(400, (20000 * 1000) >> lg))
        try:
            for mode, batched in (("single", False), ("batched", True)):
                times = {}

                mode_reps = max(5, reps // batch_jobs) if batched else reps

                def t(choice):
                    if choice not in times:
                        times[choice] = _time_ntt(lg, batched, choice, bufs, mode_reps)
                    return times[choice]

                valid = set(zntt.tile_candidates(lg))
                base = _resolved(lg, batched, (-1, -1, -1))
                # the split, each with the built-in tiles of that split
                splits = [_resolved(lg, batched, (n1, -1, -1)) for n1 in (base[0] - 1, base[0], base[0] + 1)
                          if any(c[0] == n1 for c in valid)]
                cur = min(splits, key=t)
                # pass 1's tile, then pass 2's (their ranges are independent for one split)
                cur = min([c for c in valid if c[0] == cur[0] and c[2] == cur[2]], key=t)
                choice = min([c for c in valid if c[0] == cur[0] and c[1] == cur[1]], key=t)
                zntt.set_tiles(lg, batched, *choice)
                out[mode][lg] = (choice, {"%d,%d,%d" % c: v for c, v in times.items()})
        finally:
            for b in bufs:
                b.free()
    return out


def _resolved(log_n, batched, choice):
    """The effective (log_n1, log_t1, log_t2) of a choice with -1 fields."""
    from . import ntt as zntt

    zntt.set_tiles(log_n, batched, *choice)
    return zntt.get_tiles(log_n, batched)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--min-log", type=int, default=10)
    ap.add_argument("--max-log", type=int, default=22)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--options", default="precompute=auto",
                    help="msm_init options of the table tuned on (the batched form's shared table)")
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--ntt-min-log", type=int, default=10)
    ap.add_argument("--ntt-max-log", type=int, default=23)
    ap.add_argument("--ntt-batch-jobs", type=int, default=12,
                    help="jobs per size class in the batched NTT timing (the batched form's ~256/11/2 per group)")
    ap.add_argument("--ntt-only", action="store_true", help="tune the NTT tiles only; keep the MSM table")
    ap.add_argument("--skip-ntt", action="store_true")
    args = ap.parse_args()
    from . import msm as zmsm
    from . import workload

    doc = _read(args.out)
    if not args.skip_ntt:
        res = tune_ntt(range(args.ntt_min_log, args.ntt_max_log + 1), args.ntt_batch_jobs)
        doc["ntt_tiles"] = {mode: {str(k): list(v[0]) for k, v in per.items()} for mode, per in res.items()}
        doc["ntt_seconds"] = {mode: {str(k): v[1] for k, v in per.items()} for mode, per in res.items()}
        doc["ntt_note"] = ("tiles (log_n1, log_t1, log_t2) per size class: column length 2^log_n1, log2 sequences "
                           "per workgroup of pass 1 / pass 2; seconds per transform (single) or per launch of "
                           "%d jobs (batched)" % args.ntt_batch_jobs)
        print(json.dumps({"ntt_tiles": doc["ntt_tiles"]}), flush=True)
        with open(args.out, "w") as f:
            json.dump(doc, f, indent=1)
    if args.ntt_only:
        return

    rng = np.random.default_rng(7)
    n = 1 << args.max_log
    ctx, _, _, _ = workload.generate_table(rng, args.max_log, args.options, keep_ks=False)
    ds = zmsm.DeviceBuffer(32 * n).upload(workload.rand_scalars(rng, n))
    res = tune_msm(ctx, ds.ptr, range(args.min_log, args.max_log + 1), reps=args.reps)
    table = {str(k): v[0] for k, v in res.items()}
    detail = {str(k): {str(w): t for w, t in v[1].items()} for k, v in res.items()}
    doc.update({"msm_window": table, "seconds": detail, "table_options": args.options,
                "table_points": n, "note": "window 0 = the library's own choice (size-class default or the "
                                           "window-shifted table)"})
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps({"msm_window": table}), flush=True)
    ctx.close()
    ds.free()


if __name__ == "__main__":
    main()
