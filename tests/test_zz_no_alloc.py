"""zp_msm_mult_prefix_noalloc: the MSM over a prefix of a resident table with
its working memory in the library's static per-GPU scratch array (no device
allocation).  Its results must equal zp_msm_mult_prefix_device's, byte for
byte, and the oracle's closed form, on two calls in a row, the second with a
smaller input; a job larger than the array is refused with a clear error."""
import ctypes
import random
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import msm as zmsm
from hardcaml_zprize_amd import workload
from hardcaml_zprize_amd._lib import ZpError, call

pytestmark = pytest.mark.gpu
R = oracle.R_ORDER


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(ctx, n, d_scalars):
    out = np.zeros(1, dtype=zmsm.G1_PROJ_DTYPE)
    call("zp_msm_mult_prefix_noalloc", ctx.ctx, int(n), out.ctypes.data_as(ctypes.c_void_p),
         ctypes.c_void_p(int(d_scalars) if d_scalars else 0), ctypes.c_void_p(0))
    return out[0]


def test_noalloc_equals_prefix_device_twice(msm_golden):
    _say("msm noalloc: two calls, the second smaller")
    rng = np.random.default_rng(77)
    n = 1 << 14
    ks = workload.rand_scalars(rng, n, nonzero=True)
    pts = zmsm.generate_points(ks)
    # a few points without a twisted-Edwards image, so the exceptional-point sums run too
    names = {c["name"]: c for c in msm_golden["msm_cases"]}
    special = [oracle.affine_from_json(p) for p in names["unit_intermediate_infinity_1"]["points"]]
    picks = random.Random(7).sample(range(n), 40)
    for t, i in enumerate(picks):
        pts[i] = special[t % 3]
    sc = workload.rand_scalars(rng, n)
    d = zmsm.DeviceBuffer(32 * n).upload(sc)
    try:
        with zmsm.MsmContext(pts) as ctx:
            for m in (n, 1000):
                got = _noalloc(ctx, m, d.ptr)
                want = ctx.mult_prefix_device(m, d.ptr)
                assert got.tobytes() == want.tobytes(), m
                exp = oracle.msm(pts[:m], sc[:m], 1, threads=8)[0]
                assert got.tobytes() == exp.tobytes(), m
            assert not _noalloc(ctx, 0, None)["z"].any()  # empty job: infinity
    finally:
        d.free()


def test_noalloc_refuses_jobs_beyond_the_scratch_array():
    _say("msm noalloc: too large")
    rng = np.random.default_rng(78)
    n = 1 << 17
    ctx, ks, _, _ = workload.generate_table(rng, 17)
    sc = workload.rand_scalars(rng, n)
    d = zmsm.DeviceBuffer(32 * n).upload(sc)
    try:
        with pytest.raises(ZpError, match="scratch array"):
            _noalloc(ctx, n, d.ptr)
        # the array is usable again afterwards, at a size that fits
        got = _noalloc(ctx, 1 << 12, d.ptr)
        e = oracle.dot_mod_r(sc[:1 << 12], ks[:1 << 12])
        assert oracle.affine_equal(oracle.projective_to_affine(got), oracle.g1_scalar_mul(oracle.generator(), e))
    finally:
        d.free()
        ctx.close()
