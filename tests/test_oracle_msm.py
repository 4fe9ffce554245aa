"""Pins the BLS12-377 MSM oracle (CPU only).

Checks against Python big-int textbook affine arithmetic (tests/golden/,
generated by gen_golden.py), the closed form MSM = (sum s_i k_i mod r) G for
P_i = k_i G, the special points of the reference's msm_unit_tests.rs, and
brute force (digit reconstruction) for the signed-digit model.
"""
import random

import numpy as np
import pytest

import oracle

Q, R = oracle.Q, oracle.R_ORDER


def test_fq_ops_vs_bigint():
    rng = random.Random(1)
    for _ in range(300):
        a, b = rng.randrange(Q), rng.randrange(Q)
        am, bm = oracle.to_mont(a), oracle.to_mont(b)
        assert oracle.from_mont(oracle.fq_binop("orc_fq_mul", am, bm)) == a * b % Q
        assert oracle.from_mont(oracle.fq_binop("orc_fq_add", am, bm)) == (a + b) % Q
        assert oracle.from_mont(oracle.fq_binop("orc_fq_sub", am, bm)) == (a - b) % Q
        if a:
            assert oracle.from_mont(oracle.fq_inv(am)) == pow(a, Q - 2, Q)
    for a in (0, 1, Q - 1, 2**376, 2**377 - 1 - (2**377 - Q)):
        for b in (0, 1, Q - 1):
            am, bm = oracle.to_mont(a % Q), oracle.to_mont(b)
            assert oracle.from_mont(oracle.fq_binop("orc_fq_mul", am, bm)) == a * b % Q
            assert oracle.from_mont(oracle.fq_binop("orc_fq_sub", am, bm)) == (a - b) % Q


def test_generator_and_scalar_mul(msm_golden):
    g = oracle.generator()
    assert oracle.affine_equal(g, oracle.affine_from_json(msm_golden["generator"]))
    ks = []
    for case in msm_golden["scalar_mul_G"]:
        k = oracle.words_to_int(case["k"])
        ks.append(case["k"])
        got = oracle.g1_scalar_mul(g, k)
        assert oracle.affine_equal(got, oracle.affine_from_json(case["result"])), hex(k)
    many = oracle.mul_generator_many(np.array(ks, dtype=np.uint64), threads=4)
    for got, case in zip(many, msm_golden["scalar_mul_G"]):
        assert oracle.affine_equal(got, oracle.affine_from_json(case["result"]))


def _scalar_words(ints):
    return np.array([oracle.int_to_words(s, 4) for s in ints], dtype=np.uint64)


def test_msm_golden_cases(msm_golden):
    for case in msm_golden["msm_cases"]:
        pts = np.array([oracle.affine_from_json(p) for p in case["points"]], dtype=oracle.G1_AFFINE_DTYPE)
        sc = np.array(case["scalars"], dtype=np.uint64)
        out = oracle.msm(pts, sc, 1, threads=4)
        got = oracle.projective_to_affine(out[0])
        assert oracle.affine_equal(got, oracle.affine_from_json(case["result"])), case["name"]
        if not got["infinity"]:
            # z written as Montgomery 1 (R mod q), bls12_377_g1.h:438
            assert oracle.words_to_int(out[0]["z"]) == oracle.RM


def test_msm_closed_form_multi_batch():
    rng = random.Random(99)
    n, batches = 300, 3
    ks = [rng.randrange(1, R) for _ in range(n)]
    pts = oracle.mul_generator_many(_scalar_words(ks), threads=8)
    scal = [rng.randrange(R) for _ in range(n * batches)]
    out = oracle.msm(pts, _scalar_words(scal), batches, threads=8)
    g = oracle.generator()
    for b in range(batches):
        e = sum(s * k for s, k in zip(scal[b * n:(b + 1) * n], ks)) % R
        exp = oracle.g1_scalar_mul(g, e)
        assert oracle.affine_equal(oracle.projective_to_affine(out[b]), exp)


def test_digits_reconstruct_and_bounds():
    rng = random.Random(5)
    specials = [0, 1, R - 1, 2**252, 2**253 - 1, (1 << 13) - 1, 1 << 12, (1 << 12) - 1,
                int("1" * 253, 2) % R, sum(1 << (13 * i + 12) for i in range(13)) % R]
    # 819,200 scalars in the reference test (test_scalar_transformation.ml:43-65);
    # a seeded sample here plus the patterns above.
    for s in specials + [rng.randrange(R) for _ in range(3000)] + [rng.randrange(2**253) for _ in range(500)]:
        mag, neg = oracle.msm_digits(s)
        v, off = 0, 0
        for w in range(20):
            bits = 13 if w < 13 else 12
            nb = (1 << bits) - 1 if w == 19 else 1 << (bits - 1)
            if s < R:  # the top window only stays within its buckets for s < r
                assert 0 <= mag[w] <= nb
            if w == 19:
                assert neg[w] == 0
            v += (-1 if neg[w] else 1) * int(mag[w]) << off
            off += bits
        assert off == 253 and v == s


def test_precompute_special_points(msm_golden):
    """Points with no twisted-Edwards image (msm_unit_tests.rs:30-48) report 0."""
    names = {c["name"]: c for c in msm_golden["msm_cases"]}
    pts = names["unit_intermediate_infinity_1"]["points"]
    for p in pts:
        ok, _ = oracle.precompute_point(oracle.affine_from_json(p))
        assert not ok
    ok, words = oracle.precompute_point(oracle.generator())
    assert ok
    # (y-x)/2 and (y+x)/2 with t = xy satisfy the twisted-Edwards relation
    # -x^2 + y^2 = 1 + d x^2 y^2 once unpacked.
    a = oracle.from_mont(words[0:6])
    b = oracle.from_mont(words[6:12])
    t4d = oracle.from_mont(words[12:18])
    y = (a + b) % Q
    x = (b - a) % Q
    d = 0xE2DCEDFF103E7161354A88156E4B00FE66A526A0237CFE5F683497C9AFB7635D5C9307F78E160E14F2B6D4BD949004
    assert (-x * x + y * y - 1 - d * x * x * y * y) % Q == 0
    assert t4d == 4 * d * x * y % Q


def test_planted_mistake_is_caught():
    """Dropping one term must change the closed form (guards the harness)."""
    rng = random.Random(4)
    ks = [rng.randrange(1, R) for _ in range(8)]
    pts = oracle.mul_generator_many(_scalar_words(ks))
    sc = [rng.randrange(R) for _ in range(8)]
    out = oracle.msm(pts, _scalar_words(sc), 1)
    wrong = sum(s * k for s, k in zip(sc[1:], ks[1:])) % R
    assert not oracle.affine_equal(oracle.projective_to_affine(out[0]), oracle.g1_scalar_mul(oracle.generator(), wrong))


def _digits(s):
    mag, neg = oracle.msm_digits(s)
    return [-int(m) if n else int(m) for m, n in zip(mag, neg)]


def test_digit_tie_rule():
    """A window slice (plus carry) of exactly 2^(w-1) becomes -2^(w-1) with a
    carry into the next window, and one of exactly 2^w becomes 0 with a carry
    (zprize/msm_pippenger/hardcaml/test/top/utils.ml:261-272: `>=`, not `>`).
    Expected digits written out by hand from that rule."""
    z = [0] * 20
    assert _digits(1 << 12) == [-4096, 1] + z[2:]            # slice 4096 = 2^12: the tie
    assert _digits((1 << 12) - 1) == [4095] + z[1:]          # one below the tie stays positive
    assert _digits((1 << 13) - 1) == [-1, 1] + z[2:]
    # window 1 reaches the tie only through the carry: 4095 + 1
    assert _digits((((1 << 12) - 1) << 13) | (1 << 12)) == [-4096, -4096, 1] + z[3:]
    # window 1 overflows to 2^13 through the carry: digit 0, carry on
    assert _digits((((1 << 13) - 1) << 13) | ((1 << 13) - 1)) == [-1, 0, 1] + z[3:]
    # the 12-bit windows (13..18) tie at 2^11
    assert _digits(1 << (13 * 13 + 11)) == z[:13] + [-2048, 1] + z[15:]
    # the top window (offset 241) is unsigned: slice + carry as is
    top = 2389 << 241  # (r - 1) >> 241
    assert _digits(top) == z[:19] + [2389]


def test_dot_mod_r_vs_python_ints():
    """orc_dot_mod_r (the closed form behind every large MSM check) against
    Python integers, on random and extreme operands (all-ones words: the
    640-bit accumulator's top words)."""
    rng = random.Random(31)
    for n in (1, 2, 7, 1000):
        s = [rng.randrange(2**256) for _ in range(n)]
        k = [rng.randrange(2**256) for _ in range(n)]
        got = oracle.dot_mod_r(_scalar_words(s), _scalar_words(k))
        assert got == sum(a * b for a, b in zip(s, k)) % R
    ones = [2**256 - 1] * 4096
    assert oracle.dot_mod_r(_scalar_words(ones), _scalar_words(ones)) == 4096 * (2**256 - 1) ** 2 % R
    # a planted slip (one product dropped) changes the value
    s = [rng.randrange(R) for _ in range(5)]
    k = [rng.randrange(R) for _ in range(5)]
    assert oracle.dot_mod_r(_scalar_words(s), _scalar_words(k)) != sum(a * b for a, b in zip(s[1:], k[1:])) % R


def test_numpy_dot_mod_r_matches_oracle():
    """The benchmark workload's own exact dot product (float64 limb products)
    equals the oracle's and Python's."""
    import numpy as np

    from hardcaml_zprize_amd import workload

    rng = np.random.default_rng(32)
    for n in (1, 3, (1 << 20) + 17):
        s = workload.rand_scalars(rng, n)
        k = workload.rand_scalars(rng, n, nonzero=True)
        assert workload.dot_mod_r(s, k) == oracle.dot_mod_r(s, k)
    s = np.full((9, 4), 2**64 - 1, dtype=np.uint64)
    assert workload.dot_mod_r(s, s) == 9 * (2**256 - 1) ** 2 % R
