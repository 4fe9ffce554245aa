"""The headline MSM at its measured size: N = 2^26 base points P_i = k_i G
(k_i uniform in [1, r)) generated on the GPU and kept resident, fresh scalars
uniform in [0, r) (the full 253-bit range, bit 252 included), as the
benchmark draws them.  Each result must equal (sum s_i k_i mod r) G, the
closed form of zprize/msm_pippenger/test_fpga_harness/tests/msm.rs:15-42's
comparison with arkworks (the canonical affine point is unique).

Both tables the benchmark times are run: "precompute=auto" (at 2^26 a 23-bit
window-shifted table: 11 copies, one window of 2^22 shared buckets split into
2^14 sort groups) and the default options (13 windows of 20 bits); through
zp_msm_mult_device (scalars resident), zp_msm_mult (host scalars through the
pinned uploader) and the void msm_mult entry point.
"""
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import msm as zmsm
from hardcaml_zprize_amd import workload

pytestmark = pytest.mark.gpu
LOG_N = 26
N = 1 << LOG_N


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def points_2p26():
    _say("msm 2^26: generating k_i G on the GPU")
    ks = workload.rand_scalars(np.random.default_rng(2026), N, nonzero=True)
    dk = zmsm.DeviceBuffer(32 * N).upload(ks)
    dp = zmsm.DeviceBuffer(104 * N)
    zmsm.generate_points_device(dk.ptr, dp.ptr, N)
    dk.free()
    yield ks, dp
    dp.free()


@pytest.mark.parametrize("options,shifted,width", [("precompute=auto", True, 23), ("", False, 20)])
def test_msm_2p26_closed_form(points_2p26, options, shifted, width):
    ks, dp = points_2p26
    _say(f"msm 2^26 options '{options}': msm_init")
    g = oracle.generator()
    with zmsm.MsmContext((dp.ptr, N), options) as ctx:
        assert ctx.plan() == (shifted, width, 11 if shifted else 13)
        if shifted:
            assert ctx.precompute_width() == 23
        sc = workload.rand_scalars(np.random.default_rng([2026, len(options)]), N)
        assert (sc[:, 3] >> np.uint64(60)).any()  # bit 252 occurs
        d = zmsm.DeviceBuffer(32 * N).upload(sc)
        try:
            _say(f"msm 2^26 options '{options}': zp_msm_mult_device")
            dev = ctx.mult_device(d.ptr, 1)
        finally:
            d.free()
        _say(f"msm 2^26 options '{options}': zp_msm_mult (host scalars)")
        host = ctx.mult(sc, 1)
        _say(f"msm 2^26 options '{options}': msm_mult")
        abi = ctx.mult_abi(sc, 1)
    e = oracle.dot_mod_r(sc, ks)
    want = oracle.g1_scalar_mul(g, e)
    assert oracle.affine_equal(oracle.projective_to_affine(dev[0]), want)
    assert host.tobytes() == dev.tobytes()
    assert abi.tobytes() == dev.tobytes()
