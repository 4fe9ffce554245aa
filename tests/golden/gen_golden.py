"""Generate the golden fixtures under tests/golden/ with textbook big-int math.

Independent of the reference code and of the C oracle: the Goldilocks
transforms are O(n^2) DFT sums, the BLS12-377 results use affine
short-Weierstrass formulas on y^2 = x^3 + 1 over q.  Known-answer vectors that
the reference's own tests hold are copied here as data (inputs and expected
outputs), with their source file:line.

Run:  python tests/golden/gen_golden.py   (writes ntt_golden.json, msm_golden.json)
"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))

P = 2**64 - 2**32 + 1
Q = 0x1AE3A4617C510EAC63B05C06CA1493B1A22D9F300F5138F1EF3622FBA094800170B5D44300000008508C00000000001
R_ORDER = 0x12AB655E9A2CA55660B44D1E5C37B00159AA76FED00000010A11800000000001
RM = 2**384 % Q
GX = 0x8848DEFE740A67C8FC6225BF87FF5485951E2CAA9D41BB188282C8BD37CB5CD5481512FFCD394EEAB9B16EB21BE9EF
GY = 0x1914A69C5102EFF1F674F5D30AFEEC4BD7FB348CA3E52D96D182AD44FB82305C2FE3D3634A9591AFD82DE55559C8EA6


# ----------------------------------------------------------------- NTT ----
def dft(a, inverse=False):
    n = len(a)
    w = pow(7, (P - 1) // n, P)
    if inverse:
        w = pow(w, P - 2, P)
    out = [sum(a[j] * pow(w, j * k, P) for j in range(n)) % P for k in range(n)]
    if inverse:
        ninv = pow(n, P - 2, P)
        out = [x * ninv % P for x in out]
    return out


def ntt_fixtures():
    fx = {}
    # OMEGA table, zprize/ntt/host/gf.h:65-99 (data copied from the reference).
    fx["omega_table_gf_h"] = [
        1, 18446744069414584320, 281474976710656, 18446744069397807105, 17293822564807737345,
        70368744161280, 549755813888, 17870292113338400769, 13797081185216407910,
        1803076106186727246, 11353340290879379826, 455906449640507599, 17492915097719143606,
        1532612707718625687, 16207902636198568418, 17776499369601055404, 6115771955107415310,
        12380578893860276750, 9306717745644682924, 18146160046829613826, 3511170319078647661,
        17654865857378133588, 5416168637041100469, 16905767614792059275, 9713644485405565297,
        5456943929260765144, 17096174751763063430, 1213594585890690845, 6414415596519834757,
        16116352524544190054, 9123114210336311365, 4614640910117430873, 1753635133440165772]
    # KATs: libs/hardcaml_ntt/reference_code/ntt_test.cpp:22-101 ==
    # zprize/ntt/hardcaml/test/test_ntt_against_reference.ml:33-85.
    fx["kat_8pt_linear"] = {
        "input": [1, 2, 0, 0, 0, 0, 0, 0],
        "output": [0x3, 0xfffffffefe000002, 0x0002000000000001, 0xfffffdff00000202,
                   0xffffffff00000000, 0x0000000002000001, 0xfffdffff00000002, 0x000001fffffffe01]}
    fx["kat_8pt_random"] = {
        "input": [0xcef967e3e1d0860e, 0x44be7570bcd4f9df, 0xf4848ed283e858f2, 0xa3a3a47eeb6f76f6,
                  0xa12d1d0b69c4108b, 0xeb285d19459ef6c3, 0x10d812558ad9c103, 0xd19d3e319d1b6b4a],
        "output": [0x1aaadb56e555836b, 0x975bcb9d395a282f, 0x69055db04cf94815, 0x963cdab11477cc1c,
                   0xd05b70dbcf57ddad, 0xed14bc2fbdc30962, 0x6c8e69de2cabb133, 0x9c83c8e1d49cd861]}
    fx["kat_4096pt_linear_first8"] = {
        "log_n": 12,
        "output_first8": [0x3, 0xe586a3342b3bf96c, 0x0ca769003b43919f, 0x28b1a9691a680e3c,
                          0x3b1e55b017fdb2e4, 0x309d8a339a00ae6a, 0xdc13ebf6fd47c483, 0xc12decfb84bb920e]}
    # libs/hardcaml_ntt/test/test_ntt_4step.ml:19-40 (16-point, input 0..15).
    fx["kat_16pt_4step"] = {
        "input": list(range(16)),
        "output": [120, 9185100786013534200, 18444501065828136953, 9189603281834309625,
                   18444492269600899065, 9185082089752463353, 2260596040923128,
                   9189586793186428920, 18446744069414584313, 9257157276228155385,
                   18444483473373661177, 9261661979662120952, 2251799813685240,
                   9257140787580274680, 2243003586447352, 9261643283401050105]}
    # Own brute-force DFT vectors (seeded), forward and inverse.
    rng = random.Random(20240601)
    brute = []
    for log_n in (1, 2, 3, 4, 5, 6, 7, 8):
        a = [rng.randrange(P) for _ in range(1 << log_n)]
        brute.append({"log_n": log_n, "input": a, "forward": dft(a), "inverse": dft(a, True)})
    fx["brute_force"] = brute
    # Field edge vectors, libs/hardcaml_ntt/test/test_gf.ml:9-25.
    edges = list(range(10)) + [2**31 - 1, 2**31, 2**31 + 1, 2**32 - 1, 2**32, 2**32 + 1,
                               2**63 - 1, 2**63, 2**63 + 1] + [P - 10 + i for i in range(10)]
    fx["field_edges"] = {
        "values": edges,
        "add": [[(a + b) % P for b in edges] for a in edges],
        "sub": [[(a - b) % P for b in edges] for a in edges],
        "mul": [[(a * b) % P for b in edges] for a in edges]}
    return fx


# ----------------------------------------------------------------- MSM ----
def inv(x):
    return pow(x, Q - 2, Q)


def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return (y * y - x * x * x - 1) % Q == 0


def add(p1, p2):
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    x1, y1 = p1
    x2, y2 = p2
    if x1 == x2:
        if (y1 + y2) % Q == 0:
            return None
        lam = 3 * x1 * x1 * inv(2 * y1) % Q
    else:
        lam = (y2 - y1) * inv(x2 - x1) % Q
    x3 = (lam * lam - x1 - x2) % Q
    return (x3, (lam * (x1 - x3) - y1) % Q)


def neg(p):
    return None if p is None else (p[0], (-p[1]) % Q)


def mul(p, k):
    acc = None
    while k:
        if k & 1:
            acc = add(acc, p)
        p = add(p, p)
        k >>= 1
    return acc


def msm(points, scalars):
    acc = None
    for pt, s in zip(points, scalars):
        acc = add(acc, mul(pt, s))
    return acc


def mont_words(v):
    m = v * RM % Q
    return [(m >> (64 * i)) & (2**64 - 1) for i in range(6)]


def words256(v):
    return [(v >> (64 * i)) & (2**64 - 1) for i in range(4)]


def aff_json(pt):
    if pt is None:
        return {"infinity": True}
    return {"infinity": False, "x": mont_words(pt[0]), "y": mont_words(pt[1]),
            "x_plain": hex(pt[0]), "y_plain": hex(pt[1])}


def msm_case(name, points, scalars, origin):
    res = msm(points, scalars)
    return {"name": name, "origin": origin,
            "points": [aff_json(p) for p in points],
            "scalars": [words256(s) for s in scalars],
            "result": aff_json(res)}


def msm_fixtures():
    G = (GX, GY)
    assert on_curve(G)
    assert mul(G, R_ORDER) is None
    rng = random.Random(377)
    fx = {"q": hex(Q), "r": hex(R_ORDER), "R_mod_q": hex(RM)}
    fx["generator"] = aff_json(G)
    # k*G for assorted k.
    ks = [1, 2, 3, 7, 2**252, R_ORDER - 1] + [rng.randrange(1, R_ORDER) for _ in range(6)]
    fx["scalar_mul_G"] = [{"k": words256(k), "result": aff_json(mul(G, k))} for k in ks]
    cases = []
    # Random small MSMs with P_i = k_i G.
    for n in (1, 2, 5, 16):
        ks = [rng.randrange(1, R_ORDER) for _ in range(n)]
        pts = [mul(G, k) for k in ks]
        sc = [rng.randrange(R_ORDER) for _ in range(n)]
        cases.append(msm_case(f"random_{n}", pts, sc, "seeded textbook MSM"))
    # Scalars hitting digit carries: all-ones windows, 2^w - 1 patterns, r - 1.
    pts = [mul(G, k) for k in (5, 11, 13, 17)]
    sc = [2**253 - 1 - (2**253 - R_ORDER) - 0, (1 << 252) | ((1 << 13) - 1), (1 << 13) * 4096 - 1, R_ORDER - 1]
    sc = [s % R_ORDER for s in sc]
    cases.append(msm_case("carry_patterns", pts, sc, "digit carry chains"))
    # Special points of zprize/msm_pippenger/test_fpga_harness/tests/msm_unit_tests.rs:30-48.
    s1 = (0x32D756062D349E59416ECE15CCBF8E86EF0D33183465A42FE2CB65FC1664272E6BB28F0E1C7A7C9C05824AD09ADC00,
          0x6E4B66BB23EF4BEF715F597162D6662D8161CD062D6212D39392E17232444A0760B5DC479DB98123AB3887AA3CB34E)
    s2 = (0x32D756062D349E59416ECE15CCBF8E86EF0D33183465A42FE2CB65FC1664272E6BB28F0E1C7A7C9C05824AD09ADC00,
          0x13FEEDF5CA1219ED6C9A666FB3E72D4ECA17825FAC7B17C4B5FCF4E47D703B60FAAA767E862467F615D877855C34CB3)
    s3 = (Q - 1, 0)
    for pt in (s1, s2, s3):
        assert on_curve(pt)
    assert neg(s1) == s2
    cases.append(msm_case("unit_intermediate_infinity_1", [s1, s2, s3, s3], [1, 1, 1, 1],
                          "msm_unit_tests.rs:21-72"))
    cases.append(msm_case("unit_intermediate_infinity_2", [s1, s3, neg(add(s1, s3)), s3], [1, 1, 1, 1],
                          "msm_unit_tests.rs:76-131"))
    cases.append(msm_case("unit_infinity_result", [G if i % 2 == 0 else neg(G) for i in range(256)],
                          [1] * 256, "msm_unit_tests.rs:135-180"))
    cases.append(msm_case("unit_one_bad_one_good", [G, s1], [1, 2], "msm_unit_tests.rs:240-290"))
    cases.append(msm_case("trivial_inputs", [G] * 64, [1] + [0] * 63, "util.rs:79-98"))
    cases.append(msm_case("infinity_point_input", [None, G, mul(G, 3)], [5, 7, 11], "g1_affine_t.infinity"))
    cases.append(msm_case("all_zero_scalars", [G, mul(G, 9)], [0, 0], "zero digits skipped"))
    fx["msm_cases"] = cases
    return fx


def main():
    with open(os.path.join(HERE, "ntt_golden.json"), "w") as f:
        json.dump(ntt_fixtures(), f)
    with open(os.path.join(HERE, "msm_golden.json"), "w") as f:
        json.dump(msm_fixtures(), f)


if __name__ == "__main__":
    main()
