"""BLS12-377 G1 MSM on the GPU through the C ABI, bit-exact against the oracle.

Bit-exactness is on the output bytes (x R, y R, R) / z = 0, which are unique
for a given group element, so every window layout must agree.
"""
import random
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import msm as zmsm
from hardcaml_zprize_amd import workload

pytestmark = pytest.mark.gpu
Q, R = oracle.Q, oracle.R_ORDER


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _words(ints, n):
    return np.array([oracle.int_to_words(v, n) for v in ints], dtype=np.uint64)


def _scalars(rng, count):
    return _words([rng.randrange(R) for _ in range(count)], 4)


def test_fq_mul_vs_oracle():
    _say("msm fq mul")
    rng = random.Random(3)
    vals = [0, 1, 2, Q - 1, Q - 2, 2**376, (1 << 377) - 1 - ((1 << 377) - Q)] + [rng.randrange(Q) for _ in range(2000)]
    a = [rng.choice(vals) for _ in range(4000)]
    b = [rng.choice(vals) for _ in range(4000)]
    wa = np.array([oracle.to_mont(x) for x in a])
    wb = np.array([oracle.to_mont(x) for x in b])
    got = zmsm.debug_fq_mul(wa, wb)
    for i in range(len(a)):
        assert oracle.from_mont(got[i]) == a[i] * b[i] % Q
    # and bytewise against the oracle's own Montgomery multiply
    for i in range(0, len(a), 97):
        assert np.array_equal(got[i], oracle.fq_binop("orc_fq_mul", wa[i], wb[i]))


def test_digits_reference_layout_vs_oracle():
    """M3: the reference's 20-window signed digits (utils.ml:243-288), bit-exact."""
    _say("msm digits")
    rng = random.Random(8)
    ints = [0, 1, R - 1, (1 << 252), (1 << 13) - 1, 1 << 12, sum(1 << (13 * i + 12) for i in range(13)) % R]
    ints += [rng.randrange(R) for _ in range(5000)]
    got = zmsm.debug_digits(_words(ints, 4), "window=reference")
    assert got.shape == (len(ints), 20)
    for i, s in enumerate(ints):
        mag, neg = oracle.msm_digits(s)
        exp = [-int(m) if n else int(m) for m, n in zip(mag, neg)]
        assert list(got[i]) == exp, hex(s)


def test_digits_uniform_layouts_reconstruct():
    _say("msm digits uniform")
    rng = random.Random(9)
    ints = [rng.randrange(R) for _ in range(500)] + [R - 1, 0, 1]
    for c in (8, 11, 13, 16, 19):
        got = zmsm.debug_digits(_words(ints, 4), f"window={c}")
        for i, s in enumerate(ints):
            v = sum(int(d) << (c * w) for w, d in enumerate(got[i]))
            assert v == s
            assert all(abs(int(d)) <= (1 << (c - 1)) for d in got[i][:-1])


def test_generator_vs_golden(msm_golden):
    _say("msm generator")
    cases = msm_golden["scalar_mul_G"]
    ks = np.array([c["k"] for c in cases], dtype=np.uint64)
    pts = zmsm.generate_points(ks)
    for p, c in zip(pts, cases):
        assert oracle.affine_equal(p, oracle.affine_from_json(c["result"]))


def test_precompute_vs_oracle():
    """M2: twisted-Edwards image of each point vs the oracle's host precompute
    ((y-x)/2, (y+x)/2, 4dt) of bls12_377_g1.h:606-618; ours is (y-x, y+x, 2dt)."""
    _say("msm precompute")
    rng = random.Random(12)
    ks = _words([rng.randrange(1, R) for _ in range(200)], 4)
    pts = oracle.mul_generator_many(ks)
    with zmsm.MsmContext(pts) as ctx:
        mine = ctx.precomputed(0, len(pts))
    for i, p in enumerate(pts):
        ok, ref = oracle.precompute_point(p)
        assert ok
        r = [oracle.from_mont(ref[6 * c:6 * c + 6]) for c in range(3)]
        m = [oracle.from_mont(mine[i][6 * c:6 * c + 6]) for c in range(3)]
        assert m[0] == 2 * r[0] % Q and m[1] == 2 * r[1] % Q and 2 * m[2] % Q == r[2]


@pytest.mark.parametrize("use_void_abi", [False, True])
def test_golden_cases(msm_golden, use_void_abi):
    """Textbook big-int MSMs incl. the special points of msm_unit_tests.rs."""
    for case in msm_golden["msm_cases"]:
        _say(f"msm golden {case['name']}")
        pts = np.array([oracle.affine_from_json(p) for p in case["points"]], dtype=zmsm.G1_AFFINE_DTYPE)
        sc = np.array(case["scalars"], dtype=np.uint64)
        with zmsm.MsmContext(pts) as ctx:
            out = ctx.mult_abi(sc, 1) if use_void_abi else ctx.mult(sc, 1)
        got = oracle.projective_to_affine(out[0])
        assert oracle.affine_equal(got, oracle.affine_from_json(case["result"])), case["name"]
        exp = oracle.msm(pts, sc, 1, threads=4)
        assert out.tobytes() == exp.tobytes(), case["name"]


def test_unrepresentable_points_at_chunk_boundaries(msm_golden):
    """msm_unit_tests.rs:184-236: 4096 random points with special points at
    1024, 2047, 3071, 4095, 4 batches."""
    _say("msm chunk boundaries")
    names = {c["name"]: c for c in msm_golden["msm_cases"]}
    special = [oracle.affine_from_json(p) for p in names["unit_intermediate_infinity_1"]["points"]]
    rng = random.Random(21)
    n, batches = 4096, 4
    ks = _words([rng.randrange(1, R) for _ in range(n)], 4)
    pts = zmsm.generate_points(ks)
    pts[1024] = special[0]
    pts[2047] = special[2]
    pts[3 * 1024 - 1] = special[1]
    pts[4095] = special[2]
    sc = _scalars(rng, n * batches)
    with zmsm.MsmContext(pts) as ctx:
        assert ctx.num_nonconvertible() == 4
        out = ctx.mult(sc, batches)
    exp = oracle.msm(pts, sc, batches, threads=8)
    assert out.tobytes() == exp.tobytes()


@pytest.mark.parametrize("log_n", [0, 1, 5, 10, 12])
def test_random_vs_oracle(log_n):
    _say(f"msm random log_n={log_n}")
    rng = random.Random(100 + log_n)
    n = 1 << log_n
    ks = _words([rng.randrange(1, R) for _ in range(n)], 4)
    pts = zmsm.generate_points(ks)
    sc = _scalars(rng, 2 * n)
    with zmsm.MsmContext(pts) as ctx:
        out = ctx.mult(sc, 2)
    exp = oracle.msm(pts, sc, 2, threads=8)
    assert out.tobytes() == exp.tobytes()


@pytest.mark.parametrize("options", ["window=reference", "window=8", "window=13", "window=16", "window=19",
                                     "window=12,log_u=0,log_g=0", "window=10,log_u=8,log_g=6"])
def test_layouts_agree(options):
    _say(f"msm layout {options}")
    rng = random.Random(7)
    n = 3000
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    s = [rng.randrange(R) for _ in range(n)]
    with zmsm.MsmContext(pts, options) as ctx:
        out = ctx.mult(_words(s, 4), 1)
    e = sum(a * b for a, b in zip(s, ks)) % R
    exp = oracle.g1_scalar_mul(oracle.generator(), e)
    assert oracle.affine_equal(oracle.projective_to_affine(out[0]), exp)


def test_edge_scalars():
    _say("msm edge scalars")
    rng = random.Random(31)
    n = 257
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    g = oracle.generator()
    with zmsm.MsmContext(pts) as ctx:
        zero = ctx.mult(np.zeros((n, 4), dtype=np.uint64), 1)
        assert not zero[0]["z"].any()  # infinity: z = 0
        same = [R - 1] * n  # every bucket hit in the same place (hot bucket)
        out = ctx.mult(_words(same, 4), 1)
        e = sum((R - 1) * k for k in ks) % R
        assert oracle.affine_equal(oracle.projective_to_affine(out[0]), oracle.g1_scalar_mul(g, e))
        ones = ctx.mult(np.tile(np.array([1, 0, 0, 0], dtype=np.uint64), (n, 1)), 1)
        e = sum(ks) % R
        assert oracle.affine_equal(oracle.projective_to_affine(ones[0]), oracle.g1_scalar_mul(g, e))


def test_device_scalars_and_prefix():
    _say("msm device scalars + prefix")
    rng = random.Random(44)
    n = 5000
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    s = [rng.randrange(R) for _ in range(2 * n)]
    ds = zmsm.DeviceBuffer(32 * 2 * n).upload(_words(s, 4))
    g = oracle.generator()
    with zmsm.MsmContext(pts) as ctx:
        out = ctx.mult_device(ds.ptr, 2)
        for b in range(2):
            e = sum(a * k for a, k in zip(s[b * n:(b + 1) * n], ks)) % R
            assert oracle.affine_equal(oracle.projective_to_affine(out[b]), oracle.g1_scalar_mul(g, e))
        for m in (1, 100, 4097):
            p = ctx.mult_prefix_device(m, ds.ptr)
            e = sum(a * k for a, k in zip(s[:m], ks[:m])) % R
            assert oracle.affine_equal(oracle.projective_to_affine(p), oracle.g1_scalar_mul(g, e))
    ds.free()


@pytest.mark.parametrize("log_n,options", [(20, ""), (24, ""), (24, "window=20"), (23, "window=19")])
def test_large_closed_form(log_n, options):
    """Points generated on the GPU; result vs (sum s_i k_i mod r) G.  2^20 at
    the automatic window takes the single-level sort, 2^24 (window 18) and
    the wider windows the two-level sort (window 19: a 6-bit top window)."""
    _say(f"msm large 2^{log_n} {options}")
    rng = np.random.default_rng(2024 + log_n)
    n = 1 << log_n
    ks = workload.rand_scalars(rng, n, nonzero=True)  # uniform in [1, r)
    sc = workload.rand_scalars(rng, n)  # uniform in [0, r), bit 252 included
    dk = zmsm.DeviceBuffer(32 * n).upload(ks)
    dp = zmsm.DeviceBuffer(104 * n)
    zmsm.generate_points_device(dk.ptr, dp.ptr, n)
    with zmsm.MsmContext((dp.ptr, n), options) as ctx:
        out = ctx.mult(sc, 1)
    e = oracle.dot_mod_r(sc, ks)
    assert oracle.affine_equal(oracle.projective_to_affine(out[0]), oracle.g1_scalar_mul(oracle.generator(), e))
    dk.free()
    dp.free()


def test_batch_mixed_jobs_two_tables():
    """Batched form: jobs of mixed sizes over two resident tables in one
    launch sequence; every result equals the job run alone."""
    _say("msm batch mixed")
    rng = random.Random(77)
    g = oracle.generator()
    tables = []
    for t in range(2):
        n = 3000 + 1000 * t
        ks = [rng.randrange(1, R) for _ in range(n)]
        tables.append((ks, zmsm.MsmContext(zmsm.generate_points(_words(ks, 4)))))
    jobs, expect, bufs = [], [], []
    for j in range(9):
        ks, ctx = tables[j % 2]
        n = rng.choice([0, 1, 17, 256, 1000, len(ks)])
        s = [rng.randrange(R) for _ in range(max(n, 1))]
        b = zmsm.DeviceBuffer(32 * max(n, 1)).upload(_words(s, 4))
        bufs.append(b)
        jobs.append((ctx, n, b.ptr))
        e = sum(a * k for a, k in zip(s[:n], ks[:n])) % R
        expect.append(oracle.g1_scalar_mul(g, e))
    ws = zmsm.MsmWorkspace()
    try:
        # the same jobs through each bucket sort (the two-level one also at a
        # window no single-level LDS histogram holds)
        for rep, opts in enumerate(["", "", "sort=msd", "sort=msd,window=18", "sort=atomic"]):
            out = ws.run(jobs, opts)
            for j, (o, e) in enumerate(zip(out, expect)):
                assert oracle.affine_equal(oracle.projective_to_affine(o), e), (rep, opts, j)
        out = ws.run(jobs)
        alone = [ctx.mult_prefix_device(n, p) for ctx, n, p in jobs]
        assert all(a.tobytes() == b.tobytes() for a, b in zip(alone, out))
    finally:
        ws.close()
        for b in bufs:
            b.free()
        for _, ctx in tables:
            ctx.close()


def test_many_exceptional_points(msm_golden):
    """A table that repeats the curve's points without a twisted-Edwards image
    thousands of times (grouped on the device: one exact scalar sum per group),
    interleaved with generator multiples and infinity; several batches, both
    entry points, and prefix jobs that cut through the repeated points.  The
    oracle restates the reference's per-point host fallback (driver.cpp:349-372)."""
    _say("msm many exceptional points")
    names = {c["name"]: c for c in msm_golden["msm_cases"]}
    special = [oracle.affine_from_json(p) for p in names["unit_intermediate_infinity_1"]["points"]]
    rng = random.Random(33)
    n, batches = 6000, 3
    pts = zmsm.generate_points(_words([rng.randrange(1, R) for _ in range(n)], 4))
    picks = rng.sample(range(n), 3000)
    for t, i in enumerate(picks):
        pts[i] = special[t % 3]
    pts[5] = np.zeros(1, dtype=zmsm.G1_AFFINE_DTYPE)[0]
    pts[5]["infinity"] = 1
    sc = _scalars(rng, n * batches)
    sc[picks[0:300]] = _words([R - 1] * 300, 4)  # large group sums (carries into the fifth word)
    with zmsm.MsmContext(pts) as ctx:
        assert ctx.num_nonconvertible() == 3000
        out = ctx.mult(sc, batches)
        exp = oracle.msm(pts, sc, batches, threads=8)
        assert out.tobytes() == exp.tobytes()
        assert ctx.mult_abi(sc, batches).tobytes() == exp.tobytes()
        d = zmsm.DeviceBuffer(sc.nbytes).upload(sc)
        try:
            for m in (1, 7, 1500, 4321, n):
                got = ctx.mult_prefix_device(m, d.ptr)
                want = oracle.msm(pts[:m], sc[:m], 1, threads=8)[0]
                assert got.tobytes() == want.tobytes(), m
        finally:
            d.free()


def test_hot_buckets_closed_form():
    """Scalars that put every point of a window into one bucket (all equal, all
    ones, half zero): the bucket's unit partials are summed by the log-depth
    tree, not one thread; results equal ((sum s_i k_i) mod r) G."""
    import time

    rng = random.Random(404)
    n = 1 << 20
    ks_int = [rng.randrange(1, R) for _ in range(n)]
    ks = _words(ks_int, 4)
    g = oracle.generator()
    s_eq = rng.randrange(1, R)
    cases = {
        "all_equal": np.tile(_words([s_eq], 4), (n, 1)),
        "all_ones": np.tile(np.array([[1, 0, 0, 0]], dtype=np.uint64), (n, 1)),
        "half_zero": np.concatenate([np.zeros((n // 2, 4), dtype=np.uint64),
                                     np.tile(_words([R - 1], 4), (n - n // 2, 1))]),
    }
    with zmsm.MsmContext(zmsm.generate_points(ks)) as ctx:
        for name, sc in cases.items():
            _say(f"msm hot buckets {name}")
            t0 = time.perf_counter()
            out = ctx.mult(sc, 1)
            dt = time.perf_counter() - t0
            e = oracle.dot_mod_r(sc, ks)
            assert oracle.affine_equal(oracle.projective_to_affine(out[0]), oracle.g1_scalar_mul(g, e)), name
            assert dt < 5.0, (name, dt)


@pytest.mark.parametrize("window", [8, 12, 15, 16, 17, 18, 19, 20, 21])
def test_two_level_sort_windows(window):
    """The two-level bucket sort (sort=msd) at every window width, including
    widths whose top window holds 1-6 bits (18, 19: almost every point lands
    in its few buckets): closed form ((sum s_i k_i) mod r) G, and the same
    point as the single-level sort where that one applies."""
    _say(f"msm two-level sort window={window}")
    rng = random.Random(600 + window)
    n = 5000
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    s = [rng.randrange(R) for _ in range(n)]
    sc = _words(s, 4)
    exp = oracle.g1_scalar_mul(oracle.generator(), sum(a * b for a, b in zip(s, ks)) % R)
    outs = {}
    for sort in ("msd", "atomic") + (("lds",) if window <= 16 else ()):
        with zmsm.MsmContext(pts, f"window={window},sort={sort}") as ctx:
            outs[sort] = ctx.mult(sc, 1)
        assert oracle.affine_equal(oracle.projective_to_affine(outs[sort][0]), exp), sort


@pytest.mark.parametrize("lbits", [0, 3, 7, 12])
def test_two_level_sort_split(lbits):
    """Pass B sorting 2^lbits buckets per group, from one bucket per group
    (pass A does the whole sort) to more buckets than a window has."""
    _say(f"msm two-level sort lbits={lbits}")
    rng = random.Random(700 + lbits)
    n = 3000
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    s = [rng.randrange(R) for _ in range(n)]
    with zmsm.MsmContext(pts, f"window=11,sort=msd,sort_lbits={lbits}") as ctx:
        out = ctx.mult(_words(s, 4), 1)
    exp = oracle.g1_scalar_mul(oracle.generator(), sum(a * b for a, b in zip(s, ks)) % R)
    assert oracle.affine_equal(oracle.projective_to_affine(out[0]), exp)


def test_two_level_sort_hot_buckets():
    """Hot buckets through the two-level sort at a wide window: every point of
    a window in one group and one bucket (all equal), a bucket per window at
    value 1 (all ones), and half the scalars zero."""
    import time

    rng = random.Random(405)
    n = 1 << 18
    ks_int = [rng.randrange(1, R) for _ in range(n)]
    ks = _words(ks_int, 4)
    g = oracle.generator()
    s_eq = rng.randrange(1, R)
    cases = {
        "all_equal": np.tile(_words([s_eq], 4), (n, 1)),
        "all_ones": np.tile(np.array([[1, 0, 0, 0]], dtype=np.uint64), (n, 1)),
        "half_zero": np.concatenate([np.zeros((n // 2, 4), dtype=np.uint64),
                                     np.tile(_words([R - 1], 4), (n - n // 2, 1))]),
    }
    with zmsm.MsmContext(zmsm.generate_points(ks), "window=20,sort=msd") as ctx:
        for name, sc in cases.items():
            _say(f"msm two-level hot buckets {name}")
            t0 = time.perf_counter()
            out = ctx.mult(sc, 1)
            dt = time.perf_counter() - t0
            e = oracle.dot_mod_r(sc, ks)
            assert oracle.affine_equal(oracle.projective_to_affine(out[0]), oracle.g1_scalar_mul(g, e)), name
            assert dt < 5.0, (name, dt)


@pytest.mark.parametrize("options", ["window=16,sort=msd", "window=16,sort=msd,log_u=4", "window=13,sort=atomic,log_u=8"])
def test_slices_across_many_buckets(options):
    """Scalars 1..n (window 0 holds about one entry per bucket, the higher
    windows are empty) and n copies of a few values (a handful of full
    buckets between long runs of empty ones): an accumulation slice then
    crosses hundreds of bucket ends, or lies inside one bucket; closed form
    ((sum s_i k_i) mod r) G both times."""
    rng = random.Random(810)
    n = 1 << 16
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    few = [rng.randrange(R) for _ in range(5)]
    cases = {"consecutive": [i + 1 for i in range(n)], "five_values": [few[i % 5] for i in range(n)]}
    with zmsm.MsmContext(pts, options) as ctx:
        for name, s in cases.items():
            _say(f"msm slices {options} {name}")
            out = ctx.mult(_words(s, 4), 1)
            exp = oracle.g1_scalar_mul(oracle.generator(), sum(a * b for a, b in zip(s, ks)) % R)
            assert oracle.affine_equal(oracle.projective_to_affine(out[0]), exp), name


def test_sort_option_rejected():
    from hardcaml_zprize_amd._lib import ZpError

    with pytest.raises(ZpError, match="sort"):
        zmsm.MsmContext(zmsm.generate_points(_words([1, 2, 3], 4)), "sort=radix")


def test_host_scalars_pinned_upload_batches():
    """msm_mult with host scalars large enough (64 MiB per batch) for the
    pinned multi-lane upload, over three batches (each next batch uploads
    while the previous computes), from a misaligned host array: every batch
    equals ((sum s_i k_i) mod r) G."""
    _say("msm pinned upload 3 x 2^21")
    rng = np.random.default_rng(77)
    n = 1 << 21
    ks = workload.rand_scalars(rng, n, nonzero=True)
    raw = np.zeros(3 * n * 32 + 8, dtype=np.uint8)
    sc = raw[8:].view(np.uint64).reshape(3 * n, 4)  # 8-byte offset from the allocation
    sc[:] = workload.rand_scalars(rng, 3 * n)
    dk = zmsm.DeviceBuffer(32 * n).upload(ks)
    dp = zmsm.DeviceBuffer(104 * n)
    zmsm.generate_points_device(dk.ptr, dp.ptr, n)
    try:
        with zmsm.MsmContext((dp.ptr, n)) as ctx:
            out = ctx.mult(sc, 3)
            again = ctx.mult(sc[n:2 * n], 1)
        g = oracle.generator()
        for b in range(3):
            e = oracle.dot_mod_r(sc[b * n:(b + 1) * n], ks)
            assert oracle.affine_equal(oracle.projective_to_affine(out[b]), oracle.g1_scalar_mul(g, e)), b
        assert again[0].tobytes() == out[1].tobytes()
    finally:
        dk.free()
        dp.free()


def _order2_points():
    """The three curve points (x, 0) with x^3 = -1 (order 2), as wire-form affine records."""
    w = next(pow(b, (Q - 1) // 3, Q) for b in range(2, 100) if pow(b, (Q - 1) // 3, Q) != 1)
    pts = np.zeros(3, dtype=zmsm.G1_AFFINE_DTYPE)
    for i, x in enumerate((Q - 1, (-w) % Q, (-w * w) % Q)):
        assert (x ** 3 + 1) % Q == 0
        pts[i]["x"] = oracle.to_mont(x)
        pts[i]["y"] = oracle.to_mont(0)
    return pts


def test_all_five_exceptional_points_device_vs_host_finalize(msm_golden):
    """All five curve points without a twisted-Edwards image (the three of
    order 2, the two of order 4 from msm_unit_tests.rs:30-48) repeated in a
    table, with scalar sums covering every residue mod 4: the device
    finalisation (S mod 4 multiples) equals the host finalisation (full
    double-and-add over the exact sums, option finalize=host) and the oracle
    (the reference's per-point fallback restated), for several jobs in one batch."""
    _say("msm five exceptional points, device vs host finalisation")
    names = {c["name"]: c for c in msm_golden["msm_cases"]}
    s1, s2 = [oracle.affine_from_json(p) for p in names["unit_intermediate_infinity_1"]["points"][:2]]
    special = list(_order2_points()) + [s1, s2]
    rng = random.Random(55)
    n = 2000
    pts = zmsm.generate_points(_words([rng.randrange(1, R) for _ in range(n)], 4))
    picks = rng.sample(range(n), 500)
    for t, i in enumerate(picks):
        pts[i] = special[t % 5]
    ws = zmsm.MsmWorkspace()
    bufs = []
    try:
        with zmsm.MsmContext(pts) as ctx:
            assert ctx.num_nonconvertible() == 500
            jobs, want = [], []
            for j in range(8):
                sc = _scalars(rng, n)
                if j < 4:  # force the sum of each group's scalars to residue j mod 4
                    for g in range(5):
                        idx = [i for t, i in enumerate(picks) if t % 5 == g]
                        sc[idx] = 0
                        sc[idx[0]] = oracle.int_to_words(j, 4)
                m = n if j % 2 == 0 else rng.randrange(1, n)
                b = zmsm.DeviceBuffer(32 * n).upload(sc)
                bufs.append(b)
                jobs.append((ctx, m, b.ptr))
                want.append(oracle.msm(pts[:m], sc[:m], 1, threads=8)[0])
            dev = ws.run(jobs)
            host = ws.run(jobs, "finalize=host")
            for j in range(len(jobs)):
                assert dev[j].tobytes() == want[j].tobytes(), j
                assert host[j].tobytes() == want[j].tobytes(), j
    finally:
        ws.close()
        for b in bufs:
            b.free()


def test_off_curve_point_without_image_rejected():
    """A point (x, 1) with Montgomery u = -1 (x = alpha - 1/s) is not on the
    curve: it has no twisted-Edwards image and 4P (by the group-law formulas)
    is not the identity, so msm_init rejects it (the finalisation relies on
    S P = (S mod 4) P for such points).  (A y = 0 point doubles to the
    identity under the formulas whether or not it lies on the curve, so it
    is taken as of order 2.)"""
    from hardcaml_zprize_amd._lib import ZpError

    s = 0x19D47D415B5FF60A87A8B7BBAB25EB6427DD58CA38E47030EFD1E6310AC7BF3079221BF2B4BD72C5106E9E70FCC6156
    x = (-1 - pow(s, Q - 2, Q)) % Q
    assert (1 - x ** 3 - 1) % Q != 0
    pts = zmsm.generate_points(_words([3, 5], 4))
    bad = np.zeros(1, dtype=zmsm.G1_AFFINE_DTYPE)
    bad[0]["x"] = oracle.to_mont(x)
    bad[0]["y"] = oracle.to_mont(1)
    with pytest.raises(ZpError, match="not a curve point"):
        zmsm.MsmContext(np.concatenate([pts, bad]))


def test_finalize_option_rejected():
    from hardcaml_zprize_amd._lib import ZpError

    ws = zmsm.MsmWorkspace()
    try:
        with zmsm.MsmContext(zmsm.generate_points(_words([1, 2], 4))) as ctx:
            with pytest.raises(ZpError, match="finalize"):
                ws.run([(ctx, 0, None)], "finalize=cpu")
    finally:
        ws.close()
