"""The batched form as ONE HIP graph (zp_mixed_batch): NTT groups and the
MSM launch sequence captured once and replayed.  Every replay is checked
against the oracle, with the program's own eager work (re-uploads, reads,
plain calls) between replays.  The last test runs the batched workload at
its measured size: 256 NTTs with log2 n ~ U[10, 20] (forward or inverse)
and 64 MSMs with log2 N ~ U[12, 22] over one shared 2^22-point table."""
import random
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import msm as zmsm
from hardcaml_zprize_amd import ntt as zntt
from hardcaml_zprize_amd import workload
from hardcaml_zprize_amd._lib import ZpError
from hardcaml_zprize_amd.batch import MixedBatch

pytestmark = pytest.mark.gpu
R = oracle.R_ORDER


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _words(ints, n):
    return np.array([oracle.int_to_words(v, n) for v in ints], dtype=np.uint64)


def _expected_ntts(inputs, logs, dirs):
    """Oracle transforms, grouped by (log n, direction) for the multi-threaded oracle."""
    out = [None] * len(inputs)
    groups = {}
    for i, (k, d) in enumerate(zip(logs, dirs)):
        groups.setdefault((k, d), []).append(i)
    for (k, d), idx in groups.items():
        x = np.concatenate([inputs[i] for i in idx])
        oracle.ntt_many_inplace(x, k, len(idx), inverse=bool(d), threads=8)
        for t, i in enumerate(idx):
            out[i] = x[t << k:(t + 1) << k]
    return out


def test_mixed_batch_replays_vs_oracle():
    _say("mixed batch graph, small jobs, 3 replays")
    rng = random.Random(61)
    nrng = np.random.default_rng(61)
    g = oracle.generator()
    tables = []
    for t in range(2):
        ks = [rng.randrange(1, R) for _ in range(3000 + 500 * t)]
        tables.append((ks, zmsm.MsmContext(zmsm.generate_points(_words(ks, 4)), "precompute=11" if t else "")))
    ntt_logs = [1, 3, 9, 10, 12, 12, 13, 16, 21]
    ntt_dirs = [0, 1, 0, 1, 0, 1, 0, 1, 0]
    ntt_in = [workload.rand_ntt_input(nrng, 1 << k) for k in ntt_logs]
    ntt_bufs = [zmsm.DeviceBuffer(8 << k) for k in ntt_logs]
    want_ntt = _expected_ntts(ntt_in, ntt_logs, ntt_dirs)
    msm, want_msm, bufs = [], [], []
    for j in range(10):
        ks, ctx = tables[j % 2]
        n = rng.choice([0, 1, 100, 2048, len(ks)])
        s = [rng.randrange(R) for _ in range(max(n, 1))]
        b = zmsm.DeviceBuffer(32 * max(n, 1)).upload(_words(s, 4))
        bufs.append(b)
        msm.append((ctx, n, b.ptr, 9 if j == 3 else 0))
        want_msm.append(oracle.g1_scalar_mul(g, sum(a * k for a, k in zip(s[:n], ks[:n])) % R))
    try:
        with MixedBatch([(b.ptr, k, d) for b, k, d in zip(ntt_bufs, ntt_logs, ntt_dirs)], msm, tuned=False) as mb:
            for rep in range(3):
                for b, x in zip(ntt_bufs, ntt_in):  # eager work between replays
                    b.upload(x)
                out = mb.run()
                for j, (o, w) in enumerate(zip(out, want_msm)):
                    assert oracle.affine_equal(oracle.projective_to_affine(o), w), (rep, j)
                for i, (b, w) in enumerate(zip(ntt_bufs, want_ntt)):
                    assert np.array_equal(b.download(np.empty(len(w), dtype=np.uint64)), w), (rep, i)
                alone = tables[0][1].mult_prefix_device(msm[0][1], msm[0][2])
                assert alone.tobytes() == out[0].tobytes()
    finally:
        for b in ntt_bufs + bufs:
            b.free()
        for _, ctx in tables:
            ctx.close()


def test_mixed_batch_only_ntt_or_only_msm():
    _say("mixed batch graph, one kind only")
    nrng = np.random.default_rng(62)
    x = workload.rand_ntt_input(nrng, 1 << 14)
    b = zmsm.DeviceBuffer(8 << 14).upload(x)
    try:
        with MixedBatch([(b.ptr, 14, 0)], [], tuned=False) as mb:
            assert len(mb.run()) == 0
        assert np.array_equal(b.download(np.empty(1 << 14, dtype=np.uint64)), oracle.ntt(x))
    finally:
        b.free()
    ks = workload.rand_scalars(nrng, 500, nonzero=True)
    sc = workload.rand_scalars(nrng, 500)
    d = zmsm.DeviceBuffer(32 * 500).upload(sc)
    try:
        with zmsm.MsmContext(zmsm.generate_points(ks)) as ctx, MixedBatch([], [(ctx, 500, d.ptr)], tuned=False) as mb:
            out = mb.run()
            want = oracle.g1_scalar_mul(oracle.generator(), oracle.dot_mod_r(sc, ks))
            assert oracle.affine_equal(oracle.projective_to_affine(out[0]), want)
    finally:
        d.free()


def test_mixed_batch_rejects_host_finalize():
    with zmsm.MsmContext(zmsm.generate_points(_words([1, 2], 4))) as ctx:
        with pytest.raises(ZpError, match="finalize"):
            MixedBatch([], [(ctx, 0, None)], options="finalize=host", tuned=False)


def test_batched_workload_at_size():
    """The batched workload itself (seeded as the benchmark draws it): every
    NTT against the oracle, every MSM against (sum s_i k_i mod r) G, over two
    graph replays with re-uploaded inputs, and equal to the plain calls."""
    _say("batched workload: 256 NTTs + 64 MSMs over a 2^22 table")
    rng, ntt_logs, ntt_dirs, msm_logs = workload.batched_jobs(4242)
    ctx, ks, _, _ = workload.generate_table(rng, 22, "precompute=auto")
    ntt_in = [workload.rand_ntt_input(rng, 1 << k) for k in ntt_logs]
    ntt_bufs = [zmsm.DeviceBuffer(8 << k) for k in ntt_logs]
    msm_sc = [workload.rand_scalars(rng, 1 << k) for k in msm_logs]
    msm_bufs = [zmsm.DeviceBuffer(32 << k).upload(sc) for k, sc in zip(msm_logs, msm_sc)]
    _say("batched workload: oracle transforms")
    want_ntt = _expected_ntts(ntt_in, ntt_logs, ntt_dirs)
    g = oracle.generator()
    want_msm = [oracle.g1_scalar_mul(g, oracle.dot_mod_r(sc, ks[:len(sc)])) for sc in msm_sc]
    plans = {}
    try:
        with MixedBatch([(b.ptr, k, d) for b, k, d in zip(ntt_bufs, ntt_logs, ntt_dirs)],
                        [(ctx, 1 << k, b.ptr) for k, b in zip(msm_logs, msm_bufs)]) as mb:
            for rep in range(2):
                _say(f"batched workload: graph replay {rep}")
                for b, x in zip(ntt_bufs, ntt_in):
                    b.upload(x)
                out = mb.run()
                for j, (o, w) in enumerate(zip(out, want_msm)):
                    assert oracle.affine_equal(oracle.projective_to_affine(o), w), (rep, j, msm_logs[j])
                for i, (b, w) in enumerate(zip(ntt_bufs, want_ntt)):
                    assert np.array_equal(b.download(np.empty(len(w), dtype=np.uint64)), w), (rep, i, ntt_logs[i])
            _say("batched workload: plain calls")
            for b, x in zip(ntt_bufs, ntt_in):
                b.upload(x)
            for i, (b, k, d) in enumerate(zip(ntt_bufs, ntt_logs, ntt_dirs)):
                if k not in plans:
                    plans[k] = zntt.NttGpuDriver(k, num_buffers=1)
                plans[k].inplace_device(b.ptr, d)
                assert np.array_equal(b.download(np.empty(1 << k, dtype=np.uint64)), want_ntt[i])
            for j, (k, b) in enumerate(zip(msm_logs, msm_bufs)):
                assert ctx.mult_prefix_device(1 << k, b.ptr).tobytes() == out[j].tobytes(), j
    finally:
        for p in plans.values():
            p.close()
        for b in ntt_bufs + msm_bufs:
            b.free()
        ctx.close()
