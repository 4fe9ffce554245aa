"""GPU MSM on harness test-data directories (hardcaml_zprize_amd/arkdata.py):
points.bin / scalars.bin / arkworks_results.bin as TEST_LOAD_DATA_FROM reads
them (zprize/msm_pippenger/test_fpga_harness/src/util.rs:150-173), with the
points handed to msm_init in canonical form ("coords=canonical")."""
import numpy as np
import pytest

from hardcaml_zprize_amd import arkdata as ad

import oracle

pytestmark = pytest.mark.gpu

GX = 0x008848DEFE740A67C8FC6225BF87FF5485951E2CAA9D41BB188282C8BD37CB5CD5481512FFCD394EEAB9B16EB21BE9EF
GY = 0x01914A69C5102EFF1F674F5D30AFEEC4BD7FB348CA3E52D96D182AD44FB82305C2FE3D3634A9591AFD82DE55559C8EA6

# msm_unit_tests.rs:29-47: points the twisted-Edwards map cannot take (host fallback).
BAD = [
    (0x32D756062D349E59416ECE15CCBF8E86EF0D33183465A42FE2CB65FC1664272E6BB28F0E1C7A7C9C05824AD09ADC00,
     0x6E4B66BB23EF4BEF715F597162D6662D8161CD062D6212D39392E17232444A0760B5DC479DB98123AB3887AA3CB34E),
    (0x32D756062D349E59416ECE15CCBF8E86EF0D33183465A42FE2CB65FC1664272E6BB28F0E1C7A7C9C05824AD09ADC00,
     0x13FEEDF5CA1219ED6C9A666FB3E72D4ECA17825FAC7B17C4B5FCF4E47D703B60FAAA767E862467F615D877855C34CB3),
    (0x1AE3A4617C510EAC63B05C06CA1493B1A22D9F300F5138F1EF3622FBA094800170B5D44300000008508C00000000000, 0),
]


def words(v, n):
    return np.frombuffer(v.to_bytes(8 * n, "little"), dtype="<u8")


def test_generated_directory_round_trip(tmp_path):
    """write_generated -> files -> verify; the written results also equal the
    closed form ((sum s_i k_i) R_r mod r) G from the oracle."""
    ks, scalars, results = ad.write_generated(str(tmp_path), log_n=9, batches=2, seed=5)
    n = len(ks)
    for b in range(2):
        e = oracle.dot_mod_r(scalars[b * n:(b + 1) * n], ks) * ad.R_FR % ad.R
        assert results[b] == ad.scalar_mul(e, (GX, GY))
    assert all(ad.verify(str(tmp_path), log=lambda *_: None))
    # a flipped result must be reported as a mismatch
    bad = list(results)
    bad[1] = ad.scalar_mul(2, bad[1])
    ad.write_results(str(tmp_path / "arkworks_results.bin"), bad)
    assert ad.verify(str(tmp_path), log=lambda *_: None) == [True, False]


def test_directory_with_fallback_points_and_infinity(tmp_path):
    """The harness's unit-test points that have no twisted-Edwards image, the
    point at infinity and generator multiples, in one canonical-coordinate file;
    expected results from host big-integer arithmetic."""
    G = (GX, GY)
    pts = BAD + [None, G, ad.scalar_mul(3, G), ad.scalar_mul(ad.R - 1, G)] + BAD[2:]
    n = len(pts)
    x = np.stack([words(0 if p is None else p[0], 6) for p in pts])
    y = np.stack([words(1 if p is None else p[1], 6) for p in pts])
    inf = np.array([p is None for p in pts])
    rng = np.random.default_rng(8)
    batches = 3
    svals = [int(a) << 190 ^ int(b) for a, b in rng.integers(0, 2**62, size=(batches * n, 2))]
    svals = [v % ad.R for v in svals]
    svals[0] = 0
    svals[n] = ad.R - 1
    scalars = np.stack([words(v, 4) for v in svals])
    expected = []
    for b in range(batches):
        acc = None
        for i, p in enumerate(pts):
            if p is not None:
                acc = ad._add(acc, ad.scalar_mul(svals[b * n + i] * ad.R_FR % ad.R, p))
        expected.append(acc)
    ad.write_points(tmp_path / "points.bin", x, y, inf)
    ad.write_scalars(tmp_path / "scalars.bin", scalars)
    ad.write_results(tmp_path / "arkworks_results.bin", expected)
    assert all(ad.verify(str(tmp_path), log=lambda *_: None))
    # same points through the default (Montgomery wire) coordinates agree
    from hardcaml_zprize_amd.msm import G1_AFFINE_DTYPE, MsmContext

    wire = np.zeros(n, dtype=G1_AFFINE_DTYPE)
    for i, p in enumerate(pts):
        if p is None:
            wire[i]["infinity"] = 1
        else:
            wire[i]["x"] = words(p[0] * ad.R_FQ % ad.Q, 6)
            wire[i]["y"] = words(p[1] * ad.R_FQ % ad.Q, 6)
    with MsmContext(wire) as ctx:
        assert ctx.num_nonconvertible() == 4
        assert ad.harness_msm(ctx, scalars, batches) == expected


def test_bad_coords_option_rejected():
    from hardcaml_zprize_amd._lib import ZpError
    from hardcaml_zprize_amd.msm import G1_AFFINE_DTYPE, MsmContext

    with pytest.raises(ZpError, match="coords"):
        MsmContext(np.zeros(1, dtype=G1_AFFINE_DTYPE), "coords=projective")
