"""ark-serialize test-data files of the MSM harness (hardcaml_zprize_amd/arkdata.py),
CPU side: byte layout, round trips, malformed files, and the harness's scalar
semantics (the integers it multiplies by are s R_r mod r, util.rs:115-118)."""
import os

import numpy as np
import pytest

from hardcaml_zprize_amd import arkdata as ad

import oracle

# G1 generator, canonical (bls12_377_g1.h / test_ark_bls12_377_g1.ml constants).
GX = 0x008848DEFE740A67C8FC6225BF87FF5485951E2CAA9D41BB188282C8BD37CB5CD5481512FFCD394EEAB9B16EB21BE9EF
GY = 0x01914A69C5102EFF1F674F5D30AFEEC4BD7FB348CA3E52D96D182AD44FB82305C2FE3D3634A9591AFD82DE55559C8EA6


def words(v, n):
    return np.frombuffer(v.to_bytes(8 * n, "little"), dtype="<u8")


def test_generator_and_infinity_bytes(tmp_path):
    p = tmp_path / "points.bin"
    x = np.stack([words(GX, 6), words(0, 6)])
    y = np.stack([words(GY, 6), words(1, 6)])
    ad.write_points(p, x, y, [False, True])
    raw = p.read_bytes()
    assert len(raw) == 8 + 2 * 96
    assert raw[:8] == (2).to_bytes(8, "little")
    assert raw[8:56] == GX.to_bytes(48, "little")
    assert raw[56:104] == GY.to_bytes(48, "little")          # finite point: no flag bits
    assert raw[104:152] == bytes(48)
    assert raw[152:200] == (1 | 0x40 << 376).to_bytes(48, "little")  # y = 1, infinity flag (bit 6 of byte 47)
    xr, yr, inf = ad.read_points(p)
    assert list(inf) == [False, True]
    assert ad.read_results(p) == [(GX, GY), None]
    assert (GY * GY - GX**3 - 1) % ad.Q == 0


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 257
    xs = [int(v) % ad.Q for v in rng.integers(0, 2**63, size=n)]
    x = np.stack([words((v << 300) % ad.Q, 6) for v in xs])
    y = np.stack([words((v * 7919 << 200) % ad.Q, 6) for v in xs])
    inf = rng.random(n) < 0.1
    ad.write_points(tmp_path / "p.bin", x, y, inf)
    x2, y2, inf2 = ad.read_points(tmp_path / "p.bin")
    assert (x2 == x).all() and (y2 == y).all() and (inf2 == inf).all()
    s = rng.integers(0, 2**64, size=(3 * n, 4), dtype=np.uint64)
    s[:, 3] &= np.uint64((1 << 60) - 1)
    ad.write_scalars(tmp_path / "s.bin", s)
    assert (tmp_path / "s.bin").stat().st_size == 8 + 32 * 3 * n
    assert (ad.read_scalars(tmp_path / "s.bin") == s).all()
    pts = [(GX, GY), None, ad.scalar_mul(5, (GX, GY))]
    ad.write_results(tmp_path / "r.bin", pts)
    assert ad.read_results(tmp_path / "r.bin") == pts


def test_malformed_files(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(b"\x01\x00")
    with pytest.raises(ad.FormatError):
        ad.read_points(p)
    p.write_bytes((2).to_bytes(8, "little") + bytes(96))  # says 2, holds 1
    with pytest.raises(ad.FormatError):
        ad.read_points(p)
    rec = bytearray(bytes(95) + b"\xc0")  # infinity and positive-y together
    p.write_bytes((1).to_bytes(8, "little") + bytes(rec))
    with pytest.raises(ad.FormatError):
        ad.read_points(p)
    d = tmp_path / "dir"
    d.mkdir()
    ad.write_points(d / "points.bin", np.zeros((2, 6), "<u8"), np.zeros((2, 6), "<u8"), [True, True])
    ad.write_scalars(d / "scalars.bin", np.zeros((3, 4), "<u8"))  # 3 scalars for 2 points
    ad.write_results(d / "arkworks_results.bin", [None])
    with pytest.raises(ad.FormatError):
        ad.load(d)


def test_host_group_law_against_oracle():
    """scalar_mul / wire_to_affine against the oracle's generator multiples."""
    ks = np.zeros((4, 4), dtype=np.uint64)
    vals = [1, 2, 0xDEADBEEF12345, ad.R - 1]
    for i, v in enumerate(vals):
        ks[i] = words(v, 4)
    ref = oracle.mul_generator_many(ks)
    for v, rec in zip(vals, ref):
        want = (oracle.from_mont(rec["x"]), oracle.from_mont(rec["y"]))
        assert ad.scalar_mul(v, (GX, GY)) == want
    assert ad.scalar_mul(ad.R, (GX, GY)) is None


def test_harness_scalar_semantics():
    """sum (s_i R_r mod r) P_i == R_r * sum s_i P_i on small inputs (the identity
    harness_msm relies on), with P_i = k_i G."""
    rng = np.random.default_rng(11)
    ks = [int(v) for v in rng.integers(1, 2**62, size=5)]
    ss = [int(v) * 0x9E3779B97F4A7C15 % ad.R for v in rng.integers(1, 2**62, size=5)]
    G = (GX, GY)
    lhs = None
    for s, k in zip(ss, ks):
        lhs = ad._add(lhs, ad.scalar_mul(s * ad.R_FR % ad.R, ad.scalar_mul(k, G)))
    acc = None
    for s, k in zip(ss, ks):
        acc = ad._add(acc, ad.scalar_mul(s, ad.scalar_mul(k, G)))
    assert ad.scalar_mul(ad.R_FR, acc) == lhs
    assert lhs == ad.scalar_mul(sum(s * k for s, k in zip(ss, ks)) * ad.R_FR % ad.R, G)
