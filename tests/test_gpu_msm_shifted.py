"""MSM over a window-shifted table (msm_init option "precompute"): the context
also keeps 2^(c w) P_i for every window w of a c-bit layout, and a job then
adds every window's digits into ONE bucket set.  Results must be the same
bytes as the per-window path and the oracle (zprize/msm_pippenger/host/
driver/driver.cpp:306-375 restated), including the exceptional points the
reference routes to its host fallback (driver.cpp:349-372).
"""
import random
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import msm as zmsm
from hardcaml_zprize_amd import workload

pytestmark = pytest.mark.gpu
R = oracle.R_ORDER


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _words(ints, n):
    return np.array([oracle.int_to_words(v, n) for v in ints], dtype=np.uint64)


@pytest.mark.parametrize("options,log_n", [("precompute=10", 12), ("precompute=auto", 12),
                                           ("precompute=13,window=9", 12), ("precompute=17", 16)])
def test_shifted_table_vs_oracle(options, log_n):
    """Two batches; widths 10 (26 windows), 13 (a 6-bit top window) and 17
    (2^16 shared buckets: the two-level sort splits them into groups) against
    the oracle."""
    _say(f"msm shifted {options} 2^{log_n}")
    rng = random.Random(900)
    n = 1 << log_n
    pts = zmsm.generate_points(_words([rng.randrange(1, R) for _ in range(n)], 4))
    sc = _words([rng.randrange(R) for _ in range(2 * n)], 4)
    exp = oracle.msm(pts, sc, 2, threads=8)
    with zmsm.MsmContext(pts, options) as ctx:
        out = ctx.mult(sc, 2)
    assert out.tobytes() == exp.tobytes()


def test_shifted_table_edge_scalars():
    _say("msm shifted edge scalars")
    rng = random.Random(901)
    n = 3000
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    g = oracle.generator()
    with zmsm.MsmContext(pts, "precompute=auto") as ctx:
        zero = ctx.mult(np.zeros((n, 4), dtype=np.uint64), 1)
        assert not zero[0]["z"].any()
        for s in ([R - 1] * n, [1] * n, [(1 << 252) + 12345] * n):  # every entry in one bucket per window
            out = ctx.mult(_words(s, 4), 1)
            e = sum(a * k for a, k in zip(s, ks)) % R
            assert oracle.affine_equal(oracle.projective_to_affine(out[0]), oracle.g1_scalar_mul(g, e))


def test_shifted_table_exceptional_points(msm_golden):
    """Points without a twisted-Edwards image and infinity inside the table:
    their digits are skipped in every window and the host fallback adds them."""
    _say("msm shifted exceptional points")
    names = {c["name"]: c for c in msm_golden["msm_cases"]}
    special = [oracle.affine_from_json(p) for p in names["unit_intermediate_infinity_1"]["points"]]
    rng = random.Random(902)
    n = 5000
    pts = zmsm.generate_points(_words([rng.randrange(1, R) for _ in range(n)], 4))
    picks = rng.sample(range(n), 600)
    for t, i in enumerate(picks):
        pts[i] = special[t % 3]
    pts[11] = np.zeros(1, dtype=zmsm.G1_AFFINE_DTYPE)[0]
    pts[11]["infinity"] = 1
    sc = _words([rng.randrange(R) for _ in range(2 * n)], 4)
    with zmsm.MsmContext(pts, "precompute=11") as ctx:
        out = ctx.mult(sc, 2)
    assert out.tobytes() == oracle.msm(pts, sc, 2, threads=8).tobytes()


def test_shifted_table_prefixes_and_batch():
    """Prefix jobs of a shifted table mixed with jobs on a plain table in one
    batch; each result equals the job alone, the plain table's result and the
    closed form."""
    _say("msm shifted prefixes + batch")
    rng = random.Random(903)
    g = oracle.generator()
    n = 20000
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = zmsm.generate_points(_words(ks, 4))
    s = [rng.randrange(R) for _ in range(n)]
    d = zmsm.DeviceBuffer(32 * n).upload(_words(s, 4))
    shifted = zmsm.MsmContext(pts, "precompute=12")
    plain = zmsm.MsmContext(pts)
    ws = zmsm.MsmWorkspace()
    try:
        jobs, expect = [], []
        for m in (1, 5, 300, 4096, 12345, n):
            e = sum(a * k for a, k in zip(s[:m], ks[:m])) % R
            want = oracle.g1_scalar_mul(g, e)
            for ctx in (shifted, plain):
                jobs.append((ctx, m, d.ptr))
                expect.append(want)
            got = shifted.mult_prefix_device(m, d.ptr)
            assert oracle.affine_equal(oracle.projective_to_affine(got), want), m
            assert got.tobytes() == plain.mult_prefix_device(m, d.ptr).tobytes(), m
        for opts in ("", "sort=lds"):  # the shifted jobs force the two-level sort
            out = ws.run(jobs, opts)
            for j, (o, e) in enumerate(zip(out, expect)):
                assert oracle.affine_equal(oracle.projective_to_affine(o), e), (opts, j)
    finally:
        ws.close()
        shifted.close()
        plain.close()
        d.free()


@pytest.mark.parametrize("log_n,options,width", [(20, "precompute=auto", 17), (22, "precompute=23", 23)])
def test_shifted_table_large_closed_form(log_n, options, width):
    """Points generated on the GPU, scalars uniform in [0, r) (bit 252
    included); result vs (sum s_i k_i mod r) G.  2^20 at the automatic width
    (17 bits, 15 shifted copies, 2^16 shared buckets); 2^22 over a 23-bit
    table (11 copies, 2^22 shared buckets in 2^11 sort groups), which the
    cost model takes over the per-window layout at this size."""
    _say(f"msm shifted 2^{log_n} {options} closed form")
    rng = np.random.default_rng(904 + log_n)
    n = 1 << log_n
    ks = workload.rand_scalars(rng, n, nonzero=True)
    sc = workload.rand_scalars(rng, n)
    dk = zmsm.DeviceBuffer(32 * n).upload(ks)
    dp = zmsm.DeviceBuffer(104 * n)
    zmsm.generate_points_device(dk.ptr, dp.ptr, n)
    try:
        with zmsm.MsmContext((dp.ptr, n), options) as ctx:
            assert ctx.plan() == (True, width, (253 + width - 1) // width)
            out = ctx.mult(sc, 1)
        e = oracle.dot_mod_r(sc, ks)
        assert oracle.affine_equal(oracle.projective_to_affine(out[0]), oracle.g1_scalar_mul(oracle.generator(), e))
    finally:
        dk.free()
        dp.free()


@pytest.mark.parametrize("bad", ["precompute=7", "precompute=25", "precompute=yes"])
def test_shifted_option_rejected(bad):
    with pytest.raises(RuntimeError, match="precompute"):
        zmsm.MsmContext(zmsm.generate_points(_words([1, 2], 4)), bad)
