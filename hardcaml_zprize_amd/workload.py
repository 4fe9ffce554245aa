"""Seeded synthetic inputs of the measured workloads, shared by bench.py and the tests.

Generation rules:
  - NTT vectors: uint64 uniform in [0, p), p = 2^64 - 2^32 + 1 (numpy's
    bounded integers are exactly uniform; the reference's test_random.cpp:51
    uses rng() % MODULUS, which is not).
  - MSM scalars: uniform 253-bit integers below r by rejection sampling, as
    plain little-endian words (uint64 (n, 4)); the full range, so bit 252 and
    the top window's upper half occur (about 13 % of draws have bit 252 set).
  - MSM base points: P_i = k_i G with k_i uniform in [1, r), generated on the
    GPU by the library's fixed-base comb; the k_i are kept so that a result
    can be checked against (sum s_i k_i mod r) G.
  - The batched form: 256 NTTs (log2 n ~ U[10, 20], forward/inverse ~ 1/2 each)
    and 64 MSMs (log2 N ~ U[12, 22]) over one shared 2^22-point table.
  - The served form: 70 % MSM jobs (log2 N ~ U[14, 24], one of three 2^24
    tables, a random slice of one scalar pool) and 30 % NTT jobs (log2 n ~
    U[12, 24], forward/inverse), Poisson arrivals.
"""
import time

import numpy as np

P = 2**64 - 2**32 + 1
R_ORDER = 0x12AB655E9A2CA55660B44D1E5C37B00159AA76FED00000010A11800000000001
R_WORDS = tuple((R_ORDER >> (64 * i)) & (2**64 - 1) for i in range(4))


def rand_ntt_input(rng, n):
    """n uint64 uniform in [0, p)."""
    return rng.integers(0, P, size=n, dtype=np.uint64)


def rand_scalars(rng, n, nonzero=False):
    """n uniform 253-bit integers below r (rejection sampling), uint64 (n, 4).
    nonzero: uniform in [1, r) instead."""
    out = np.empty((n, 4), dtype=np.uint64)
    filled = 0
    r = [np.uint64(w) for w in R_WORDS]
    while filled < n:
        m = int((n - filled) * 1.8) + 64
        c = rng.integers(0, 2**64, size=(m, 4), dtype=np.uint64)
        c[:, 3] &= np.uint64((1 << 61) - 1)
        lt = (c[:, 3] < r[3]) | ((c[:, 3] == r[3]) & ((c[:, 2] < r[2]) | ((c[:, 2] == r[2]) & (
            (c[:, 1] < r[1]) | ((c[:, 1] == r[1]) & (c[:, 0] < r[0]))))))
        if nonzero:
            lt &= c.any(axis=1)
        c = c[lt]
        take = min(len(c), n - filled)
        out[filled:filled + take] = c[:take]
        filled += take
    return out


def generate_table(rng, log_n, options="", keep_ks=True):
    """A resident MSM table of 2^log_n points P_i = k_i G, generated on the GPU.

    Returns (ctx, ks, gen_s, init_s); ks is None unless keep_ks."""
    from . import msm as zmsm
    from ._lib import call

    n = 1 << log_n
    ks = rand_scalars(rng, n, nonzero=True)
    dk = zmsm.DeviceBuffer(32 * n).upload(ks)
    dp = zmsm.DeviceBuffer(104 * n)
    try:
        t0 = time.perf_counter()
        zmsm.generate_points_device(dk.ptr, dp.ptr, n)
        call("zp_device_synchronize")
        t1 = time.perf_counter()
        ctx = zmsm.MsmContext((dp.ptr, n), options)
        t2 = time.perf_counter()
    finally:
        dk.free()
        dp.free()
    return ctx, (ks if keep_ks else None), t1 - t0, t2 - t1


def dot_mod_r(s, k):
    """sum_i s_i k_i mod r for uint64 (n, 4) arrays of integers < 2^256.

    Exact: both operands are split into 16-bit limbs held in float64, and
    the limb-pair sums are formed by matrix products over blocks of 2^20
    rows (each block sum < 2^52, so every partial sum is an exact double);
    the 256 block sums are then combined with Python integers."""
    s = np.ascontiguousarray(s, dtype=np.uint64).reshape(-1, 4)
    k = np.ascontiguousarray(k, dtype=np.uint64).reshape(-1, 4)
    assert s.shape == k.shape
    acc = [[0] * 16 for _ in range(16)]
    block = 1 << 20
    for b0 in range(0, len(s), block):
        sl = s[b0:b0 + block].view(np.uint16).astype(np.float64)  # (m, 16) little-endian limbs
        kl = k[b0:b0 + block].view(np.uint16).astype(np.float64)
        m = sl.T @ kl  # (16, 16)
        mi = m.astype(np.uint64)
        for a in range(16):
            row = acc[a]
            for b in range(16):
                row[b] += int(mi[a, b])
    total = 0
    for a in range(16):
        for b in range(16):
            total += acc[a][b] << (16 * (a + b))
    return total % R_ORDER


def batched_jobs(seed):
    """The batched form's jobs: (ntt_logs, ntt_dirs, msm_logs) from one seed."""
    rng = np.random.default_rng(seed)
    ntt_logs = [int(x) for x in rng.integers(10, 21, size=256)]
    ntt_dirs = [int(x) for x in rng.integers(0, 2, size=256)]
    msm_logs = [int(x) for x in rng.integers(12, 23, size=64)]
    return rng, ntt_logs, ntt_dirs, msm_logs


def served_specs(rng, njobs, tables=3, table_log=24):
    """Job descriptions of the served stream (dicts; MSM: log_n, table,
    offset into the scalar pool; NTT: log_n, direction)."""
    specs = []
    for _ in range(njobs):
        if rng.random() < 0.7:
            k = int(rng.integers(14, 25))
            off = int(rng.integers(0, (1 << table_log) - (1 << k) + 1))
            specs.append(dict(kind="msm", log_n=k, table=int(rng.integers(0, tables)), off=off))
        else:
            k = int(rng.integers(12, 25))
            specs.append(dict(kind="ntt", log_n=k, direction=int(rng.integers(0, 2))))
    return specs
