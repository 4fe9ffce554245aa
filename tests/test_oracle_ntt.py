"""Pins the Goldilocks/NTT oracle (CPU only).

Checks against: the reference's KAT vectors and OMEGA table (fixtures), the
reference's own gf.h compiled into oracle/_ref (when built), Python big ints,
O(n^2) brute-force DFTs and closed forms.
"""
import random

import numpy as np
import pytest

import oracle

P = oracle.P


def test_field_edges_vs_bigint(ntt_golden):
    fe = ntt_golden["field_edges"]
    vals = fe["values"]
    for i, a in enumerate(vals):
        for j, b in enumerate(vals):
            a_ = a % P
            b_ = b % P
            assert oracle.gl_add(a_, b_) == fe["add"][i][j] % P
            assert oracle.gl_sub(a_, b_) == fe["sub"][i][j] % P
            assert oracle.gl_mul(a_, b_) == fe["mul"][i][j] % P


def test_field_random_vs_bigint():
    rng = random.Random(7)
    for _ in range(20000):
        a, b = rng.randrange(P), rng.randrange(P)
        assert oracle.gl_mul(a, b) == a * b % P
        assert oracle.gl_add(a, b) == (a + b) % P
        assert oracle.gl_sub(a, b) == (a - b) % P


def test_field_vs_reference_gf_h():
    ref = oracle.ref_gf()
    if ref is None:
        pytest.skip("oracle/_ref/libgf_ref.so not built (reference checkout absent)")
    rng = np.random.default_rng(11)
    a = rng.integers(0, P, size=200000, dtype=np.uint64)
    b = rng.integers(0, P, size=200000, dtype=np.uint64)
    edges = np.array([0, 1, 2, P - 1, P - 2, 2**32 - 1, 2**32, 2**63, 2**64 - 2**32], dtype=np.uint64)
    a = np.concatenate([a, np.repeat(edges, len(edges))])
    b = np.concatenate([b, np.tile(edges, len(edges))])
    out = np.zeros_like(a)
    ref.ref_gf_mul_many(oracle._ptr(a), oracle._ptr(b), oracle._ptr(out), len(a))
    exp = np.array([(int(x) * int(y)) % P for x, y in zip(a[:2000], b[:2000])], dtype=np.uint64)
    assert np.array_equal(out[:2000], exp)
    for x, y, r in zip(a[::97], b[::97], out[::97]):
        assert oracle.gl_mul(int(x), int(y)) == int(r)
        assert oracle.gl_add(int(x), int(y)) == ref.ref_gf_add(int(x), int(y))
        assert oracle.gl_sub(int(x), int(y)) == ref.ref_gf_sub(int(x), int(y))
    for i in range(33):
        assert oracle.gl_root(i) == ref.ref_gf_omega(i)


def test_roots_vs_omega_table(ntt_golden):
    tab = ntt_golden["omega_table_gf_h"]
    for i, w in enumerate(tab):
        assert oracle.gl_root(i) == w
        if i > 0:
            assert pow(w, 2**i, P) == 1 and pow(w, 2 ** (i - 1), P) == P - 1


@pytest.mark.parametrize("name", ["kat_8pt_linear", "kat_8pt_random"])
def test_kat_8pt(ntt_golden, name):
    kat = ntt_golden[name]
    assert list(oracle.ntt(kat["input"])) == kat["output"]


def test_kat_4096_linear(ntt_golden):
    kat = ntt_golden["kat_4096pt_linear_first8"]
    a = np.zeros(4096, dtype=np.uint64)
    a[0], a[1] = 1, 2
    out = oracle.ntt(a)
    assert list(out[:8]) == kat["output_first8"]
    # closed form 1 + 2 w^k for every k
    w = oracle.gl_root(12)
    for k in range(0, 4096, 37):
        assert int(out[k]) == (1 + 2 * pow(w, k, P)) % P


def test_kat_16pt_four_step(ntt_golden):
    kat = ntt_golden["kat_16pt_4step"]
    assert list(oracle.ntt(kat["input"])) == kat["output"]
    assert list(oracle.ntt_four_step(kat["input"], 2)) == kat["output"]


def test_brute_force(ntt_golden):
    for case in ntt_golden["brute_force"]:
        assert list(oracle.ntt(case["input"])) == case["forward"]
        assert list(oracle.ntt(case["input"], inverse=True)) == case["inverse"]


def test_four_step_equals_direct_64x64():
    # the 2^12 workload: n = 4096 as a 64 x 64 4-step on the software model
    rng = np.random.default_rng(0)
    a = rng.integers(0, P, size=4096, dtype=np.uint64)
    assert np.array_equal(oracle.ntt_four_step(a, 6), oracle.ntt(a))
    for log_rows in (1, 3, 9):
        assert np.array_equal(oracle.ntt_four_step(a, log_rows), oracle.ntt(a))


def test_closed_forms_and_roundtrip():
    for log_n in (1, 5, 10, 13):
        n = 1 << log_n
        delta = np.zeros(n, dtype=np.uint64)
        delta[0] = 1
        assert np.all(oracle.ntt(delta) == 1)
        ones = np.ones(n, dtype=np.uint64)
        out = oracle.ntt(ones)
        assert int(out[0]) == n % P and np.all(out[1:] == 0)
        rng = np.random.default_rng(log_n)
        a = rng.integers(0, P, size=n, dtype=np.uint64)
        assert np.array_equal(oracle.ntt(oracle.ntt(a), inverse=True), a)
        # unscaled inverse = n * scaled inverse
        un = oracle.ntt(a, inverse=True, scaled=False)
        sc = oracle.ntt(a, inverse=True)
        assert all(int(u) == int(s) * n % P for u, s in zip(un[:50], sc[:50]))


def test_linearity():
    rng = np.random.default_rng(3)
    n = 1 << 9
    a = rng.integers(0, P, size=n, dtype=np.uint64)
    b = rng.integers(0, P, size=n, dtype=np.uint64)
    s = np.array([(int(x) + int(y)) % P for x, y in zip(a, b)], dtype=np.uint64)
    fa, fb, fs = oracle.ntt(a), oracle.ntt(b), oracle.ntt(s)
    assert all(int(z) == (int(x) + int(y)) % P for x, y, z in zip(fa, fb, fs))


def test_planted_mistake_is_caught():
    """A wrong root (a transposed index) must break the closed form."""
    a = np.zeros(16, dtype=np.uint64)
    a[0], a[1] = 1, 2
    out = oracle.ntt(a)
    w_wrong = oracle.gl_root(5)
    assert any(int(out[k]) != (1 + 2 * pow(w_wrong, k, P)) % P for k in range(16))


@pytest.mark.parametrize("log_r", [1, 2, 3, 4, 5])
def test_reverse_core_closed_form(log_r):
    """orc_ntt_reverse (reference.cpp:60-84 restated step by step) against the
    element formula its steps compose to: with R = 2^log_r and w = OMEGA[2 log_r],
    out[i][j] = in[R-1-j][R-1-i] * w^(j (R-1-i))."""
    R = 1 << log_r
    rng = np.random.default_rng(log_r)
    a = rng.integers(0, P, size=R * R, dtype=np.uint64)
    got = oracle.ntt_reverse(a, log_r)
    w = pow(7, (P - 1) >> (2 * log_r), P)
    for i in range(R):
        for j in range(R):
            want = int(a[(R - 1 - j) * R + (R - 1 - i)]) * pow(w, j * (R - 1 - i), P) % P
            assert int(got[i * R + j]) == want
