"""Benchmark: 2^24 Goldilocks NTT (headline), 2^26 BLS12-377 MSM, batched and served forms.

Prints ONE JSON line.  `value` is the 2^24 forward NTT throughput in GB/s on
the algorithmic 16*n-byte basis (one compulsory read + one write of the
vector), inputs resident in HBM, timed with HIP events on the stream the
kernels run on.  The MSM, batched and served results ride along in the same
line.  One GPU only.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # 100 untimed transforms (~22 ms): the GPU reaches its steady clocks after
    # ~20 ms of load (2^24 NTT after 5 warm-up transforms 233-235 us, after
    # 100 or 400: 216 us)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--ramp-ms", type=float, default=60.0,
                    help="untimed transforms for at least this long before the warm-up steps (clock ramp)")
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--msm-log-n", type=int, default=26)
    ap.add_argument("--msm-steps", type=int, default=3)
    ap.add_argument("--msm-warmup", type=int, default=1)
    ap.add_argument("--batch-options", default="precompute=auto",
                    help="msm_init option string of the batched form's shared 2^22 table (both forms use it)")
    ap.add_argument("--msm-options", default="precompute=auto",
                    help="msm_init option string of the 2^N table ('' = default options; otherwise the "
                         "default-option table is timed too, as plain_table)")
    ap.add_argument("--serve-jobs", type=int, default=240)
    ap.add_argument("--serve-load", type=float, default=6.0)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-batch-ms", type=float, default=100.0,
                    help="estimated work per served batch (ms); 0 = every waiting job that fits the budget")
    ap.add_argument("--serve-load2", type=float, default=0.7,
                    help="second offered load (x plain capacity) for the latency comparison; 0 = skip")
    ap.add_argument("--skip-msm", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-host", action="store_true", help="skip the host-buffer (PCIe-inclusive) NTT figures")
    ap.add_argument("--host-evals", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 2^log_n u64 (PMC scaling reference) and exit")
    return ap.parse_args()


def bench_ntt(args, torch):
    from hardcaml_zprize_amd import ntt as zntt
    from hardcaml_zprize_amd import workload

    log_n = args.log_n
    n = 1 << log_n
    rng = np.random.default_rng(args.seed)
    host = workload.rand_ntt_input(rng, n)
    dev = torch.from_numpy(host.view(np.int64)).to("cuda:0")
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    drv = zntt.NttGpuDriver(log_n)
    # The first transform of the seeded input is kept: the CPU-baseline leg
    # compares it with the CPU transform of the same input.
    drv.inplace_device(dev.data_ptr(), zntt.FORWARD, sptr)
    torch.cuda.synchronize()
    first_out = dev.cpu().numpy().view(np.uint64).copy()
    # Clock ramp: the GPU reaches its steady clocks only after ~20 ms of load
    # (2^24 after 5 warm-up transforms: 233-235 us; after 100: 216 us), so
    # untimed transforms run for at least --ramp-ms before the W warm-up steps.
    t_ramp = time.perf_counter()
    ramp = 0
    while (time.perf_counter() - t_ramp) * 1e3 < args.ramp_ms or ramp < 4:
        for _ in range(16):
            drv.inplace_device(dev.data_ptr(), zntt.FORWARD, sptr)
        torch.cuda.synchronize()
        ramp += 16
    for _ in range(args.warmup):
        drv.inplace_device(dev.data_ptr(), zntt.FORWARD, sptr)
    torch.cuda.synchronize()
    # Timed region: K transforms back to back between one start and one end
    # event on the launch stream; per-transform events (each marker costs
    # ~2 us per transform) are taken in a second loop below.
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        drv.inplace_device(dev.data_ptr(), zntt.FORWARD, sptr)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = ev0.elapsed_time(ev1)
    # Per-transform events (median / min per transform, markers included).
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    evs[0].record(stream)
    for i in range(args.steps):
        drv.inplace_device(dev.data_ptr(), zntt.FORWARD, sptr)
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)]
    ms_marked = evs[0].elapsed_time(evs[-1]) / args.steps
    # Per-kernel split (same stream, HIP events around each launch) for the
    # roofline of the dominant kernel.
    ms_pass1, ms_pass2 = drv.time_passes(dev.data_ptr(), zntt.FORWARD, args.steps, sptr)
    ms = total_ms / args.steps
    alg_bytes = 16 * n
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    per_sorted = sorted(per)
    # Correctness of the timed path: the vector went through warmup+steps
    # forward transforms; check one more transform against the oracle only on
    # small sizes (the parity tests cover 2^24 bit-exact).
    res = {
        "ms_per_step": ms,
        "ms_per_step_with_event_per_transform": ms_marked,
        "gbs": gbs,
        "wall_s": wall,
        "median_ms": per_sorted[len(per) // 2],
        "min_ms": per_sorted[0],
        "alg_bytes": alg_bytes,
        "ms_pass1": ms_pass1,
        "ms_pass2": ms_pass2,
        "first_out": first_out,
        "ramp_transforms": ramp,
    }
    del dev
    drv.close()
    return res


def bench_ntt_host(args):
    """The NTT through host buffers, PCIe included: simple_evaluate latency
    (pageable numpy vectors; bench_latency.cpp's blocking flow) and the buffer
    API throughput of bench_throughput.cpp:60-107 (every buffer loaded once,
    E evaluations enqueued round-robin without waiting, then E waits), whose
    copies overlap the serialized transforms.  Every buffer's output is
    checked against simple_evaluate of its input."""
    import numpy as np

    from hardcaml_zprize_amd import ntt as zntt

    n = 1 << args.log_n
    P = 2**64 - 2**32 + 1
    rng = np.random.default_rng(args.seed + 1)
    res = {}
    with zntt.NttGpuDriver(args.log_n) as drv:
        x = rng.integers(0, P, size=n, dtype=np.uint64)
        lat = []
        for i in range(4):
            t0 = time.perf_counter()
            y = drv.simple_evaluate(x)
            lat.append(time.perf_counter() - t0)
        lat = sorted(lat[1:])
        res["simple_evaluate_ms"] = round(lat[len(lat) // 2] * 1e3, 3)
        nb = drv.max_num_buffers_in_flight()
        inputs = [rng.integers(0, P, size=n, dtype=np.uint64) for _ in range(nb)]
        want = [drv.simple_evaluate(a) for a in inputs]
        bufs = [drv.request_buffer() for _ in range(nb)]
        for b, a in zip(bufs, inputs):
            drv.transfer_input_vector_to_user_buffer(b, a)
        evals = max(args.host_evals, nb)
        t0 = time.perf_counter()
        for e in range(evals):
            drv.enqueue_evaluation_async(bufs[e % nb])
        for e in range(evals):
            drv.wait_for_result(bufs[e % nb])
        secs = time.perf_counter() - t0
        ok = all(np.array_equal(drv.transfer_user_buffer_to_output_vector(b), w) for b, w in zip(bufs, want))
        for b in bufs:
            drv.free_buffer(b)
        del y
    if not ok:
        raise RuntimeError("buffer-API throughput run: an output differs from simple_evaluate")
    res.update({
        "buffers": nb, "evaluations": evals,
        "throughput_ntts_per_s": round(evals / secs, 2),
        "throughput_gbs_16n": round(evals * 16 * n / secs / 1e9, 2),
        "basis": "host pinned buffers -> device -> host; 16*n bytes cross PCIe per transform (8n each way)",
    })
    return res


Q_FIELD = 0x1AE3A4617C510EAC63B05C06CA1493B1A22D9F300F5138F1EF3622FBA094800170B5D44300000008508C00000000001
G1_GEN = (0x8848DEFE740A67C8FC6225BF87FF5485951E2CAA9D41BB188282C8BD37CB5CD5481512FFCD394EEAB9B16EB21BE9EF,
          0x1914A69C5102EFF1F674F5D30AFEEC4BD7FB348CA3E52D96D182AD44FB82305C2FE3D3634A9591AFD82DE55559C8EA6)


def closed_form_wire(e):
    """The wire words (x R, y R, R) of e G, or None for infinity: textbook
    Jacobian double-and-add on y^2 = x^3 + 1 with Python integers.  bench.py's
    own check of the timed MSM results (P_i = k_i G gives sum s_i P_i =
    (sum s_i k_i mod r) G)."""
    q = Q_FIELD
    acc = None  # Jacobian (X, Y, Z)

    def dbl(p):
        X, Y, Z = p
        if Y == 0:
            return None
        A, B = X * X % q, Y * Y % q
        C = B * B % q
        D = 2 * ((X + B) ** 2 - A - C) % q
        E = 3 * A % q
        X3 = (E * E - 2 * D) % q
        return (X3, (E * (D - X3) - 8 * C) % q, 2 * Y * Z % q)

    def add_aff(p, x2, y2):
        if p is None:
            return (x2, y2, 1)
        X1, Y1, Z1 = p
        z2 = Z1 * Z1 % q
        U2, S2 = x2 * z2 % q, y2 * z2 * Z1 % q
        H, r = (U2 - X1) % q, (S2 - Y1) % q
        if H == 0:
            return dbl(p) if r == 0 else None
        HH = H * H % q
        HHH = H * HH % q
        V = X1 * HH % q
        X3 = (r * r - HHH - 2 * V) % q
        return (X3, (r * (V - X3) - Y1 * HHH) % q, Z1 * H % q)

    for bit in bin(e)[2:] if e else "":
        acc = dbl(acc) if acc is not None else None
        if bit == "1":
            acc = add_aff(acc, *G1_GEN)
    if acc is None:
        return None
    X, Y, Z = acc
    zi = pow(Z, q - 2, q)
    x, y = X * zi * zi % q, Y * zi * zi * zi % q
    rm = 2**384 % q
    return tuple(((v * rm % q) >> (64 * i)) & (2**64 - 1) for v in (x, y, 1) for i in range(6))


def result_words(out):
    """Wire words of one g1_projective_t record, or None when z = 0 (infinity)."""
    if not out["z"].any():
        return None
    return tuple(int(w) for f in ("x", "y", "z") for w in out[f])


def dot_mod_r_device(torch, s, k):
    """sum_i s_i k_i mod r for uint64 (n, 4) host arrays, on the GPU with
    torch int64 arithmetic over 16-bit limbs: per block of 2^22 rows, each
    of the 16 x 16 limb-pair sums is a column sum of products < 2^32, below
    2^54 (exact); blocks are accumulated as Python integers."""
    from hardcaml_zprize_amd import workload

    n = len(s)
    acc = [[0] * 16 for _ in range(16)]
    block = 1 << 22
    for b0 in range(0, n, block):
        sl = torch.from_numpy(s[b0:b0 + block].view(np.int16)).to("cuda:0").to(torch.int64) & 0xFFFF
        kl = torch.from_numpy(k[b0:b0 + block].view(np.int16)).to("cuda:0").to(torch.int64) & 0xFFFF
        sums = torch.stack([(sl[:, a:a + 1] * kl).sum(dim=0) for a in range(16)]).cpu().numpy()
        for a in range(16):
            for c in range(16):
                acc[a][c] += int(sums[a, c])
    total = 0
    for a in range(16):
        for c in range(16):
            total += acc[a][c] << (16 * (a + c))
    return total % workload.R_ORDER


def bench_msm(args, probe, torch):
    """The 2^N MSM on a table loaded with --msm-options (default: with its
    window-shifted copies, "precompute=auto"), then the same table loaded with
    the default options for comparison.  Every call's result is checked
    against (sum s_i k_i mod r) G; a mismatch raises."""
    res = bench_msm_table(args, probe, torch, args.msm_options)
    res["options"] = args.msm_options
    if args.msm_options:
        plain = bench_msm_table(args, probe, torch, "", steps=max(1, args.msm_steps - 1), multi=False)
        res["plain_table"] = {"options": "", "ms": plain["ms"], "ms_msm_mult_host_scalars":
                              plain["ms_msm_mult_host_scalars"], "init_s": plain["init_s"],
                              "verified_calls": plain["verified_calls"]}
    return res


def check_msm(torch, out, sc, ks, what):
    want = closed_form_wire(dot_mod_r_device(torch, sc, ks))
    if result_words(out) != want:
        raise RuntimeError(f"MSM result differs from (sum s_i k_i mod r) G: {what}")


def bench_msm_table(args, probe, torch, options, steps=None, multi=True):
    from hardcaml_zprize_amd import msm as zmsm
    from hardcaml_zprize_amd import workload

    log_n = args.msm_log_n
    n = 1 << log_n
    steps = args.msm_steps if steps is None else steps
    rng = np.random.default_rng(args.seed + 1000)
    log(f"MSM 2^{log_n}: generating base points on the GPU")
    ctx, ks, gen_s, init_s = workload.generate_table(rng, log_n, options)
    log(f"MSM points generated in {gen_s:.2f} s, msm_init ({options or 'default options'}) {init_s:.2f} s")
    dsc = zmsm.DeviceBuffer(32 * n)
    dev_ms, host_ms = [], []
    verified = []
    for call_i in range(args.msm_warmup + steps):
        sc = workload.rand_scalars(np.random.default_rng([args.seed, call_i]), n)  # fresh per call
        t0 = time.perf_counter()
        out_host = ctx.mult(sc, 1)  # msm_mult with host scalars (H2D included)
        t1 = time.perf_counter()
        dsc.upload(sc)
        t2 = time.perf_counter()
        out_dev = ctx.mult_device(dsc.ptr, 1)  # scalars resident in HBM
        t3 = time.perf_counter()
        if call_i >= args.msm_warmup:
            host_ms.append((t1 - t0) * 1e3)
            dev_ms.append((t3 - t2) * 1e3)
        check_msm(torch, out_dev[0], sc, ks, f"2^{log_n} call {call_i}, resident scalars, options '{options}'")
        if out_host.tobytes() != out_dev.tobytes():
            raise RuntimeError(f"msm_mult (host scalars) differs from the resident-scalar result, call {call_i}")
        verified.append(call_i)
        log(f"MSM call {call_i}: msm_mult {1e3 * (t1 - t0):.1f} ms, resident scalars {1e3 * (t3 - t2):.1f} ms, "
            f"both equal (sum s_i k_i mod r) G")
    dsc.free()
    # msm_mult over several batches (the harness's num_batches): the upload of
    # batch b + 1 overlaps the computation of batch b.
    batches = 2
    multi_ms = None
    if multi:
        sc = np.concatenate([workload.rand_scalars(np.random.default_rng([args.seed, 99, b]), n)
                             for b in range(batches)])
        t0 = time.perf_counter()
        out = ctx.mult(sc, batches)
        multi_ms = (time.perf_counter() - t0) * 1e3 / batches
        for b in range(batches):
            check_msm(torch, out[b], sc[b * n:(b + 1) * n], ks, f"2^{log_n} msm_mult batch {b} of {batches}")
        del sc
        log(f"MSM msm_mult x{batches} batches: {multi_ms:.1f} ms per batch, each equal to its closed form")
    ctx.close()
    del ks
    ms = float(np.median(dev_ms))
    work = 140 * n * 288  # reference algorithm's v_mad_u64_u32 lane-ops (20 windows x 7 Fq muls x 288 MACs)
    peak = probe["mad_u64_u32_lane_ops_per_s"]
    return {"n": n, "ms": ms, "ms_msm_mult_host_scalars": float(np.median(host_ms)),
            "ms_per_batch_msm_mult_2_batches": multi_ms, "point_gen_s": gen_s,
            "init_s": init_s, "verified": True, "verified_calls": len(verified) + (batches if multi else 0),
            "check": "every call (warm-up, timed, multi-batch) equals (sum s_i k_i mod r) G; scalars uniform "
                     "in [0, r), k_i uniform in [1, r)",
            "roofline": {"bound": "valu", "achieved": work / (ms * 1e-3), "peak": peak, "unit": "mad_u64_u32 lane-op/s",
                         "frac": work / (ms * 1e-3) / peak,
                         "basis": "140 N Fq muls x 288 v_mad_u64_u32 (reference algorithm) / resident-scalar MSM time"}}


def bench_batched(args, torch):
    """256 NTTs (log2 n ~ U[10,20]) + 64 MSMs (log2 N ~ U[12,22]) over one 2^22 table:
    plain one-at-a-time calls vs the whole batch as ONE HIP graph launch
    (zp_mixed_batch: NTT groups and the MSM launch sequence as parallel
    branches), per-job MSM windows from the committed autotuner table."""
    from hardcaml_zprize_amd import autotune
    from hardcaml_zprize_amd import msm as zmsm
    from hardcaml_zprize_amd import ntt as zntt
    from hardcaml_zprize_amd import workload
    from hardcaml_zprize_amd._lib import call
    from hardcaml_zprize_amd.batch import MixedBatch

    rng, ntt_logs, ntt_dirs, msm_logs = workload.batched_jobs(args.seed + 2000)
    ctx, ks, _, _ = workload.generate_table(rng, 22, args.batch_options)
    ntt_in = [workload.rand_ntt_input(rng, 1 << k) for k in ntt_logs]
    ntt_bufs = [zmsm.DeviceBuffer(8 << k) for k in ntt_logs]
    msm_sc = [workload.rand_scalars(rng, 1 << k) for k in msm_logs]
    msm_bufs = [zmsm.DeviceBuffer(32 << k).upload(sc) for k, sc in zip(msm_logs, msm_sc)]
    plans = {k: zntt.NttGpuDriver(k, num_buffers=1) for k in set(ntt_logs)}
    table = autotune.load()
    ntt_tiles = autotune.load_ntt_tiles()
    autotune.apply_ntt_tiles(ntt_tiles)  # both forms: single-transform tiles (plain), batch-group tiles (graph)
    plain_ws = zmsm.MsmWorkspace()
    s1 = ctypes.c_void_p()
    call("zp_stream_create", ctypes.byref(s1))

    def reset_inputs():
        for b, x in zip(ntt_bufs, ntt_in):
            b.upload(x)

    def ntt_outputs():
        return [b.download(np.empty(1 << k, dtype=np.uint64)) for b, k in zip(ntt_bufs, ntt_logs)]

    def plain():
        lat = []
        res = []
        t0 = time.perf_counter()
        for i in range(256):
            plans[ntt_logs[i]].inplace_device(ntt_bufs[i].ptr, ntt_dirs[i])
            call("zp_device_synchronize")
            lat.append(time.perf_counter() - t0)
        for i in range(64):  # one job per call, with the same tuned window the batch uses
            res.append(plain_ws.run([(ctx, 1 << msm_logs[i], msm_bufs[i].ptr, table.get(msm_logs[i], 0))])[0])
            lat.append(time.perf_counter() - t0)
        return time.perf_counter() - t0, lat, res

    t0 = time.perf_counter()
    mb = MixedBatch([(b.ptr, k, d) for b, k, d in zip(ntt_bufs, ntt_logs, ntt_dirs)],
                    [(ctx, 1 << k, b.ptr) for k, b in zip(msm_logs, msm_bufs)], table=table)
    create_s = time.perf_counter() - t0

    def batched():
        t0 = time.perf_counter()
        mb.launch(s1.value)
        res = mb.wait(s1.value)
        total = time.perf_counter() - t0
        return total, [total] * 320, res

    # correctness: every NTT output of the graph equals the plain call's, every
    # MSM result equals the plain call's and (sum s_i k_i mod r) G
    reset_inputs()
    _, _, plain_msm = plain()
    want_ntt = ntt_outputs()
    for rep in range(2):  # graph replays, eager work (re-uploads, reads) between them
        reset_inputs()
        _, _, got_msm = batched()
        got_ntt = ntt_outputs()
        if not all(np.array_equal(a, b) for a, b in zip(want_ntt, got_ntt)):
            raise RuntimeError(f"batched graph replay {rep}: an NTT differs from its one-at-a-time result")
        if not all(a.tobytes() == b.tobytes() for a, b in zip(plain_msm, got_msm)):
            raise RuntimeError(f"batched graph replay {rep}: an MSM differs from its one-at-a-time result")
    for i, k in enumerate(msm_logs):
        n = 1 << k
        if result_words(plain_msm[i]) != closed_form_wire(dot_mod_r_device(torch, msm_sc[i], ks[:n])):
            raise RuntimeError(f"batched MSM job {i} (2^{k}) differs from (sum s_i k_i mod r) G")
    reps = 3
    tp = [plain() for _ in range(reps)]
    tb = [batched() for _ in range(reps)]
    plain_t = float(np.median([t for t, _, _ in tp]))
    batch_t = float(np.median([t for t, _, _ in tb]))
    lat_p = np.array(tp[len(tp) // 2][1])
    lat_b = np.array(tb[len(tb) // 2][1])
    mb.close()
    plain_ws.close()
    for p in plans.values():
        p.close()
    for b in ntt_bufs + msm_bufs:
        b.free()
    call("zp_stream_destroy", s1)
    ctx.close()
    jobs = 320
    return {"jobs": jobs, "plain_jobs_per_s": jobs / plain_t, "batched_jobs_per_s": jobs / batch_t,
            "speedup": plain_t / batch_t, "plain_s": plain_t, "batched_s": batch_t,
            "plain_p50_ms": float(np.percentile(lat_p, 50) * 1e3), "plain_p99_ms": float(np.percentile(lat_p, 99) * 1e3),
            "batched_p50_ms": float(np.percentile(lat_b, 50) * 1e3),
            "batched_p99_ms": float(np.percentile(lat_b, 99) * 1e3),
            "graph": True, "graph_create_s": create_s,
            "autotune_table": os.path.basename(autotune.DEFAULT_PATH) if table else None,
            "tuned_windows": {str(k): v for k, v in sorted(table.items())},
            "tuned_ntt_tiles": {mode: {str(k): list(v) for k, v in sorted(per.items())}
                                for mode, per in ntt_tiles.items()},
            "table_options": args.batch_options,
            "checked": "graph replays equal the one-at-a-time results (NTT vectors and MSM bytes); every MSM "
                       "equals (sum s_i k_i mod r) G",
            "note": "inputs resident; NTTs in place; the batch's latency is its completion time; plain "
                    "latency is each job's completion time from the start of the sequence"}


def bench_served(args):
    """Poisson stream: 70% MSM (log2 N in [14,24], 3 resident 2^24 tables), 30% NTT
    (log2 n in [12,24], fwd/inv), HBM budget 200 GB; plain one-at-a-time vs continuous batching."""
    import numpy as np

    from hardcaml_zprize_amd import msm as zmsm
    from hardcaml_zprize_amd import serve

    from hardcaml_zprize_amd import workload

    rng = np.random.default_rng(args.seed + 3000)
    tables = []
    for t in range(3):
        ctx, _, _, _ = workload.generate_table(rng, 24, keep_ks=False)
        tables.append(ctx)
    P = workload.P
    pool = workload.rand_scalars(rng, 1 << 24)
    ntt_pool = rng.integers(0, P, size=1 << 24, dtype=np.uint64)
    njobs = args.serve_jobs
    specs = []
    for i in range(njobs):
        if rng.random() < 0.7:
            k = int(rng.integers(14, 25))
            off = int(rng.integers(0, (1 << 24) - (1 << k) + 1))
            specs.append(dict(kind=serve.MSM, log_n=k, table=int(rng.integers(0, 3)), off=off))
        else:
            k = int(rng.integers(12, 25))
            specs.append(dict(kind=serve.NTT, log_n=k, direction=int(rng.integers(0, 2))))

    def jobs():
        out = []
        for i, sp in enumerate(specs):
            if sp["kind"] == serve.MSM:
                n = 1 << sp["log_n"]
                out.append(serve.Job(serve.MSM, id=i, log_n=sp["log_n"], table=sp["table"],
                                     scalars=pool[sp["off"]:sp["off"] + n]))
            else:
                out.append(serve.Job(serve.NTT, id=i, log_n=sp["log_n"], direction=sp["direction"],
                                     data=ntt_pool[:1 << sp["log_n"]]))
        return out

    # calibrate: plain server with every job queued at t = 0
    plain = serve.GpuServer(tables, hbm_budget=200e9, workers=1, batching=False)
    serve.run_stream(plain, jobs()[:8], [0.0] * 8)  # warm-up
    cal_s, _ = serve.run_stream(plain, jobs(), [0.0] * njobs)
    batch_s = args.serve_batch_ms * 1e-3 if args.serve_batch_ms > 0 else float("inf")
    served = serve.GpuServer(tables, hbm_budget=200e9, workers=args.serve_workers, batching=True,
                             max_batch_s=batch_s)
    serve.run_stream(served, jobs()[:8], [0.0] * 8)  # warm-up

    def at_load(load):
        rate = load * njobs / cal_s  # offered load relative to plain capacity
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=njobs))
        arrivals -= arrivals[0]
        jp = jobs()
        plain_s, plain_lat = serve.run_stream(plain, jp, list(arrivals))
        b0 = served.batches
        js = jobs()
        served_s, served_lat = serve.run_stream(served, js, list(arrivals))
        same = all((a.result.tobytes() == b.result.tobytes()) for a, b in zip(jp, js))
        return rate, plain_s, plain_lat, served_s, served_lat, served.batches - b0, same

    rate, plain_s, plain_lat, served_s, served_lat, batches, same = at_load(args.serve_load)
    low = None
    if args.serve_load2 > 0:
        r2, p2, pl2, s2, sl2, nb2, same2 = at_load(args.serve_load2)
        same = same and same2
        low = {"offered_load_x_plain_capacity": args.serve_load2, "offered_rate_jobs_per_s": r2,
               "plain_jobs_per_s": njobs / p2, "served_jobs_per_s": njobs / s2,
               "plain_latency_ms": {k: v * 1e3 for k, v in serve.percentiles(pl2).items()},
               "served_latency_ms": {k: v * 1e3 for k, v in serve.percentiles(sl2).items()},
               "served_batches": nb2}
    plain.close()
    served.close()
    bound = device_bound(specs, tables, pool, ntt_pool)
    for t in tables:
        t.close()
    return {"jobs": njobs, "offered_rate_jobs_per_s": rate, "plain_capacity_jobs_per_s": njobs / cal_s,
            "device_only_bound_jobs_per_s": bound["jobs_per_s"],
            "served_over_device_bound": (njobs / served_s) / bound["jobs_per_s"],
            "plain_over_device_bound": (njobs / plain_s) / bound["jobs_per_s"],
            "device_only": bound,
            "plain_jobs_per_s": njobs / plain_s, "served_jobs_per_s": njobs / served_s,
            "speedup": plain_s / served_s,
            "plain_latency_ms": {k: v * 1e3 for k, v in serve.percentiles(plain_lat).items()},
            "served_latency_ms": {k: v * 1e3 for k, v in serve.percentiles(served_lat).items()},
            "served_batches": batches, "workers": args.serve_workers, "hbm_budget_gb": 200,
            "offered_load_x_plain_capacity": args.serve_load, "max_batch_ms": args.serve_batch_ms,
            "below_saturation": low,
            "results_equal_plain": bool(same),
            "note": "host inputs (H2D/D2H inside each job); MSM scalars are slices of one seeded pool"}


def device_bound(specs, tables, pool, ntt_pool):
    """Upper bound of any schedule of the served jobs on this GPU: each job
    timed alone with its inputs already in HBM (no copies), on the same
    library paths and tuned windows; jobs/s = jobs / sum of these times."""
    from hardcaml_zprize_amd import autotune
    from hardcaml_zprize_amd import msm as zmsm
    from hardcaml_zprize_amd import ntt as zntt
    from hardcaml_zprize_amd._lib import call

    windows = autotune.load()
    autotune.apply_ntt_tiles()
    d_pool = zmsm.DeviceBuffer(pool.nbytes).upload(pool)
    d_ntt = zmsm.DeviceBuffer(ntt_pool.nbytes).upload(ntt_pool)
    ws = zmsm.MsmWorkspace()
    plans = {}
    t_msm = t_ntt = 0.0
    try:
        for rep in range(2):  # the first pass grows workspaces and builds plans
            t_msm = t_ntt = 0.0
            for sp in specs:
                k = sp["log_n"]
                if sp["kind"] == "msm":
                    ptr = d_pool.ptr + 32 * sp["off"]
                    t0 = time.perf_counter()
                    ws.run([(tables[sp["table"]], 1 << k, ptr, windows.get(k, 0))])
                    t_msm += time.perf_counter() - t0
                else:
                    if k not in plans:
                        plans[k] = zntt.NttGpuDriver(k, num_buffers=1)
                    t0 = time.perf_counter()
                    plans[k].inplace_device(d_ntt.ptr, sp["direction"])
                    call("zp_device_synchronize")
                    t_ntt += time.perf_counter() - t0
    finally:
        ws.close()
        for p in plans.values():
            p.close()
        d_pool.free()
        d_ntt.free()
    total = t_msm + t_ntt
    return {"jobs_per_s": len(specs) / total, "msm_s": t_msm, "ntt_s": t_ntt,
            "basis": "every job alone, inputs resident in HBM, synchronous; sum of times"}


def probe_rates():
    from hardcaml_zprize_amd._lib import call

    res = {}
    for kind, name in ((0, "mad_u64_u32_lane_ops_per_s"), (1, "valu32_lane_ops_per_s")):
        v = ctypes.c_double()
        call("zp_probe_valu_rate", kind, ctypes.byref(v))
        res[name] = v.value
    log(f"VALU probe: {res}")
    return res


def cpu_baseline_ntt(args, gpu_first_out, budget_s=12.0):
    """The oracle's restated ntt_reference (zprize/ntt/host/reference.cpp:87-137)
    on one host thread, over the same seeded input as the GPU leg; its output
    is compared with the GPU's first transform of that input."""
    import oracle  # checker / CPU baseline only
    from hardcaml_zprize_amd import workload

    n = 1 << args.log_n
    a = workload.rand_ntt_input(np.random.default_rng(args.seed), n)
    runs = 0
    t0 = time.perf_counter()
    while True:
        out = oracle.ntt(a)
        runs += 1
        el = time.perf_counter() - t0
        if el > budget_s or runs >= 5:
            break
    per = el / runs
    same = gpu_first_out is not None and np.array_equal(out, gpu_first_out)
    if not same:
        raise RuntimeError(f"GPU 2^{args.log_n} NTT differs from the CPU reference on seed {args.seed}")
    return {"value": 16 * n / per / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{runs} x 2^{args.log_n} forward NTT, restated zprize/ntt/host/reference.cpp ntt_reference, "
                      f"1 thread, {per:.3f} s each, seed {args.seed}",
            "gpu_output_bit_exact": bool(same)}


def ntt_2p12_model(args, torch):
    """The 2^12 forward NTT as a 64 x 64 4-step on the CPU software model
    (the oracle's restated libs/hardcaml_ntt/src/reference_model.ml:92-105,
    1 thread), beside the direct CPU transform and the GPU's 2^12 transform
    of the same seeded input (seed 0)."""
    import oracle  # checker / CPU baseline only
    from hardcaml_zprize_amd import ntt as zntt
    from hardcaml_zprize_amd import workload

    n = 4096
    a = workload.rand_ntt_input(np.random.default_rng(0), n)

    def per_call(fn, budget=1.0):
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget or k < 3:
            fn()
            k += 1
        return (time.perf_counter() - t0) / k

    model = oracle.ntt_four_step(a, 6)
    direct = oracle.ntt(a)
    t_model = per_call(lambda: oracle.ntt_four_step(a, 6))
    t_direct = per_call(lambda: oracle.ntt(a))
    dev = torch.from_numpy(a.view(np.int64)).to("cuda:0")
    stream = torch.cuda.current_stream()
    with zntt.NttGpuDriver(12) as drv:
        drv.inplace_device(dev.data_ptr(), zntt.FORWARD, stream.cuda_stream)
        torch.cuda.synchronize()
        gpu = dev.cpu().numpy().view(np.uint64).copy()
        reps = 2000
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            drv.inplace_device(dev.data_ptr(), zntt.FORWARD, stream.cuda_stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        gpu_us = ev0.elapsed_time(ev1) * 1e3 / reps
    ok = np.array_equal(model, direct) and np.array_equal(gpu, direct)
    if not ok:
        raise RuntimeError("2^12 NTT: the 64x64 four-step model, the direct transform and the GPU disagree")
    return {"workload": "forward Goldilocks NTT n=2^12 (64x64 4-step), seed 0, natural order",
            "cpu_model_us": t_model * 1e6, "cpu_model_cores": 1,
            "cpu_direct_us": t_direct * 1e6, "gpu_us": gpu_us,
            "gpu_timing": f"{reps} back-to-back transforms between two HIP events, data resident",
            "bit_exact": True}


def cpu_baseline_msm(args, log_n=20):
    """The oracle's restated reference host algorithm (20 signed windows,
    twisted-Edwards buckets, running-sum reduction; driver.cpp:306-375,
    pippenger.h), windows in parallel on the host cores, on one 2^20 sample of
    the 2^26 workload (points k_i G generated on the GPU, scalars uniform in
    [0, r)); its result is compared with the GPU MSM of the same inputs."""
    import oracle  # checker / CPU baseline only
    from hardcaml_zprize_amd import msm as zmsm
    from hardcaml_zprize_amd import workload

    threads = min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(args.seed + 77)
    n = 1 << log_n
    ks = workload.rand_scalars(rng, n, nonzero=True)
    pts = zmsm.generate_points(ks)
    sc = workload.rand_scalars(rng, n)
    t0 = time.perf_counter()
    cpu = oracle.msm(pts, sc, 1, threads=threads)
    el = time.perf_counter() - t0
    with zmsm.MsmContext(pts) as ctx:
        gpu = ctx.mult(sc, 1)
    if gpu.tobytes() != cpu.tobytes():
        raise RuntimeError("2^20 MSM: GPU result differs from the CPU reference path")
    return {"value": n / el, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"one 2^{log_n}-point MSM (restated reference host Pippenger, {threads} threads) in {el:.2f} s; "
                      f"2^26 at this rate: {el * (1 << 26) / n:.1f} s",
            "ms_2p20": el * 1e3, "gpu_result_bit_exact": True}


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "ntt2p24_k_ntt4096_hbm_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch") and kernel in line:
                    return int(line.split()[-1])
    except OSError:
        pass
    return None


def split_label(log_n):
    """The 4-step split the library uses for one transform of 2^log_n points
    (built-in tiles): 4096 x 4096 at 2^24."""
    from hardcaml_zprize_amd import ntt as zntt

    if log_n >= 24:
        return "4096x4096"
    if log_n < zntt.TILE_MIN_LOG:  # up to 2^9: one workgroup, one pass
        return "single-pass"
    n1 = zntt.get_tiles(log_n, False)[0]
    return f"{1 << n1}x{1 << (log_n - n1)}"


def ntt_roofline(args, ntt):
    """HBM roofline of the 2^24 transform: 16 n algorithmic bytes (one read +
    one write of the vector) over the whole-transform device time.  Each of
    the two kernels is one half of the transform and is credited 8 n bytes
    over its own mean HIP-event time (per_kernel)."""
    n = 1 << args.log_n
    alg = 16 * n
    achieved = alg / (ntt["ms_per_step"] * 1e-3) / 1e9
    per_kernel = []
    traffic = 0
    names = ("k_ntt4096_cols", "k_ntt4096r16_rows") if args.log_n == 24 else ("pass 1", "pass 2")
    for name, kms in zip(names, (ntt["ms_pass1"], ntt["ms_pass2"])):
        a = 8 * n / (kms * 1e-3) / 1e9
        t = measured_traffic(name) if args.log_n == 24 else None
        traffic = None if (t is None or traffic is None) else traffic + t
        per_kernel.append({"kernel": name, "ms": round(kms, 5), "alg_bytes": 8 * n, "achieved": round(a, 3),
                           "frac": round(a / HBM_PEAK_GBS, 5), "traffic": t})
    return {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "traffic_over_alg": round(traffic / alg, 4) if traffic else None,
            "per_kernel": per_kernel,
            "basis": "achieved = 16 n bytes / whole-transform time (K back-to-back transforms between two HIP "
                     "events); per_kernel: 8 n bytes / that launch's mean HIP-event time on the launch stream; "
                     "traffic: rocprofv3 FETCH_SIZE + WRITE_SIZE per transform, copy-kernel scaled"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)
    if args.copy_probe:
        n = 1 << args.log_n
        src = torch.empty(n, dtype=torch.int64, device="cuda:0").fill_(7)
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.steps} copies of {8 * n} bytes")
        return
    log(f"NTT 2^{args.log_n}: warmup {args.warmup}, steps {args.steps}")
    ntt = bench_ntt(args, torch)
    log(f"NTT {ntt['ms_per_step'] * 1e3:.1f} us/transform, {ntt['gbs']:.1f} GB/s")
    out = {
        "metric": "2^24 NTT GB/s, 2^26 MSM ms; batched vs plain jobs/s + p99, bit-exact",
        "value": round(ntt["gbs"], 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ntt["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform in [0,p))",
        "config": {"workload": f"Goldilocks forward NTT n=2^{args.log_n} in place ({split_label(args.log_n)} 4-step), "
                               f"resident in HBM, seed {args.seed}", "log_n": args.log_n},
        "roofline": ntt_roofline(args, ntt),
    }
    out["ntt"] = {"median_ms": ntt["median_ms"], "min_ms": ntt["min_ms"],
                  "ms_per_step_with_event_per_transform": ntt["ms_per_step_with_event_per_transform"],
                  "timing": "value: K back-to-back transforms between two events after W warm-up "
                            "transforms; median/min: a second loop with an event after every transform"}
    if not args.skip_host:
        out["ntt_host"] = bench_ntt_host(args)
        log(f"NTT through host buffers: {out['ntt_host']}")
    out["valu_probe"] = probe = probe_rates()
    if not args.skip_msm:
        out["msm"] = bench_msm(args, probe, torch)
        log(f"MSM 2^{args.msm_log_n}: {out['msm']['ms']:.1f} ms (resident scalars)")
    if not args.skip_batch:
        out["batched"] = bench_batched(args, torch)
        log(f"batched: {out['batched']}")
    if not args.skip_serve:
        out["served"] = bench_served(args)
        log(f"served: {out['served']}")
    if not args.skip_cpu:
        log("CPU baseline (oracle, 1 thread)")
        out["cpu_baseline"] = cpu_baseline_ntt(args, ntt["first_out"])
        out["ntt_2p12_model"] = ntt_2p12_model(args, torch)
        log(f"2^12 four-step model: {out['ntt_2p12_model']}")
        if not args.skip_msm:
            out["msm"]["cpu_baseline"] = cpu_baseline_msm(args)
            log(f"MSM CPU baseline: {out['msm']['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
