"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the CPU-baseline
leg of bench.py may import this package, and only as the checker; the shipped
library (hardcaml_zprize_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libgf_ref.so")

P = 2**64 - 2**32 + 1
Q = 0x1AE3A4617C510EAC63B05C06CA1493B1A22D9F300F5138F1EF3622FBA094800170B5D44300000008508C00000000001
R_ORDER = 0x12AB655E9A2CA55660B44D1E5C37B00159AA76FED00000010A11800000000001
RM = 2**384 % Q

# g1_affine_t: x[6], y[6] u64 + bool + 7 pad = 104 bytes (rust_types.h:70-74).
G1_AFFINE_DTYPE = np.dtype([("x", "<u8", (6,)), ("y", "<u8", (6,)), ("infinity", "u1"), ("_pad", "u1", (7,))])
G1_PROJ_DTYPE = np.dtype([("x", "<u8", (6,)), ("y", "<u8", (6,)), ("z", "<u8", (6,))])
assert G1_AFFINE_DTYPE.itemsize == 104 and G1_PROJ_DTYPE.itemsize == 144

_lib = None
_ref = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64 = ctypes.c_uint64
        vp = ctypes.c_void_p
        for name in ("orc_gl_add", "orc_gl_sub", "orc_gl_mul", "orc_gl_pow"):
            getattr(L, name).restype = u64
            getattr(L, name).argtypes = [u64, u64]
        L.orc_gl_inv.restype = u64
        L.orc_gl_inv.argtypes = [u64]
        L.orc_gl_root.restype = u64
        L.orc_gl_root.argtypes = [ctypes.c_uint]
        L.orc_ntt_inplace.argtypes = [vp, ctypes.c_uint, ctypes.c_int]
        L.orc_ntt_inplace_unscaled.argtypes = [vp, ctypes.c_uint, ctypes.c_int]
        L.orc_ntt_four_step.argtypes = [vp, ctypes.c_uint, ctypes.c_uint]
        L.orc_ntt_many.argtypes = [vp, ctypes.c_uint, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
        L.orc_ntt_reverse.argtypes = [vp, vp, ctypes.c_uint]
        for name in ("orc_fq_mul", "orc_fq_add", "orc_fq_sub"):
            getattr(L, name).argtypes = [vp, vp, vp]
        for name in ("orc_fq_inv", "orc_fq_to_mont", "orc_fq_from_mont"):
            getattr(L, name).argtypes = [vp, vp]
        L.orc_msm_digits.argtypes = [vp, vp, vp]
        L.orc_msm_precompute_point.restype = ctypes.c_int
        L.orc_msm_precompute_point.argtypes = [vp, vp]
        L.orc_msm.argtypes = [vp, ctypes.c_size_t, vp, ctypes.c_size_t, vp, ctypes.c_int]
        L.orc_g1_add.argtypes = [vp, vp, vp]
        L.orc_g1_scalar_mul.argtypes = [vp, vp, vp]
        L.orc_g1_generator.argtypes = [vp]
        L.orc_g1_mul_generator_many.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.orc_g1_projective_to_affine.argtypes = [vp, vp]
        L.orc_dot_mod_r.argtypes = [vp, vp, ctypes.c_size_t, vp]
        _lib = L
    return _lib


def ref_gf():
    """The reference's own gf.h compiled into oracle/_ref, or None if absent."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        L = ctypes.CDLL(REF_LIB_PATH)
        u64 = ctypes.c_uint64
        for name in ("ref_gf_add", "ref_gf_sub", "ref_gf_mul"):
            getattr(L, name).restype = u64
            getattr(L, name).argtypes = [u64, u64]
        L.ref_gf_omega.restype = u64
        L.ref_gf_omega.argtypes = [ctypes.c_uint]
        L.ref_gf_mul_many.argtypes = [ctypes.c_void_p] * 3 + [u64]
        _ref = L
    return _ref


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- NTT ----
def gl_add(a, b):
    return lib().orc_gl_add(a, b)


def gl_sub(a, b):
    return lib().orc_gl_sub(a, b)


def gl_mul(a, b):
    return lib().orc_gl_mul(a, b)


def gl_root(log_n):
    return lib().orc_gl_root(log_n)


def ntt(a, inverse=False, scaled=True):
    """Return the transform of a uint64 vector (natural order in and out)."""
    x = np.ascontiguousarray(a, dtype=np.uint64).copy()
    log_n = int(len(x)).bit_length() - 1
    assert len(x) == 1 << log_n
    if scaled:
        lib().orc_ntt_inplace(_ptr(x), log_n, 1 if inverse else 0)
    else:
        lib().orc_ntt_inplace_unscaled(_ptr(x), log_n, 1 if inverse else 0)
    return x


def ntt_reverse(a, log_row_size):
    """The REVERSE debug core's output for a (4^log_row_size uint64)."""
    src = np.ascontiguousarray(a, dtype=np.uint64)
    assert len(src) == 1 << (2 * log_row_size)
    dst = np.empty_like(src)
    lib().orc_ntt_reverse(_ptr(dst), _ptr(src), log_row_size)
    return dst


def ntt_four_step(a, log_rows):
    x = np.ascontiguousarray(a, dtype=np.uint64).copy()
    log_n = int(len(x)).bit_length() - 1
    lib().orc_ntt_four_step(_ptr(x), log_n, log_rows)
    return x


def ntt_many_inplace(x, log_n, count, inverse=False, threads=1):
    assert x.dtype == np.uint64 and x.flags.c_contiguous and x.size >= count << log_n
    lib().orc_ntt_many(_ptr(x), log_n, count, 1 if inverse else 0, threads)


# ---------------------------------------------------------------- MSM ----
def int_to_words(v, n):
    return np.array([(v >> (64 * i)) & (2**64 - 1) for i in range(n)], dtype=np.uint64)


def words_to_int(w):
    return sum(int(x) << (64 * i) for i, x in enumerate(w))


def to_mont(v):
    return int_to_words(v * RM % Q, 6)


def from_mont(w):
    return words_to_int(w) * pow(RM, Q - 2, Q) % Q


def fq_binop(name, a, b):
    out = np.zeros(6, dtype=np.uint64)
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    getattr(lib(), name)(_ptr(out), _ptr(a), _ptr(b))
    return out


def fq_inv(a):
    out = np.zeros(6, dtype=np.uint64)
    a = np.ascontiguousarray(a, dtype=np.uint64)
    lib().orc_fq_inv(_ptr(out), _ptr(a))
    return out


def msm_digits(scalar_int):
    s = int_to_words(scalar_int, 4)
    mag = np.zeros(20, dtype=np.uint32)
    neg = np.zeros(20, dtype=np.uint8)
    lib().orc_msm_digits(_ptr(s), _ptr(mag), _ptr(neg))
    return mag, neg


def precompute_point(aff):
    """aff: one G1_AFFINE_DTYPE record. Returns (ok, 18 words)."""
    a = np.array([aff], dtype=G1_AFFINE_DTYPE)
    out = np.zeros(18, dtype=np.uint64)
    ok = lib().orc_msm_precompute_point(_ptr(a), _ptr(out))
    return bool(ok), out


def msm(points, scalars, num_batches=1, threads=8):
    """points: G1_AFFINE_DTYPE array; scalars: uint64 (num_batches*N, 4)."""
    points = np.ascontiguousarray(points, dtype=G1_AFFINE_DTYPE)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64)
    assert scalars.shape == (num_batches * len(points), 4)
    out = np.zeros(num_batches, dtype=G1_PROJ_DTYPE)
    lib().orc_msm(_ptr(points), len(points), _ptr(scalars), num_batches, _ptr(out), threads)
    return out


def generator():
    g = np.zeros(1, dtype=G1_AFFINE_DTYPE)
    lib().orc_g1_generator(_ptr(g))
    return g[0]


def mul_generator_many(ks, threads=8):
    """ks: uint64 (n, 4) scalars; returns G1_AFFINE_DTYPE array of k_i * G."""
    ks = np.ascontiguousarray(ks, dtype=np.uint64)
    out = np.zeros(len(ks), dtype=G1_AFFINE_DTYPE)
    lib().orc_g1_mul_generator_many(_ptr(out), _ptr(ks), len(ks), threads)
    return out


def dot_mod_r(s, k):
    """sum_i s_i k_i mod r for uint64 (n, 4) arrays."""
    s = np.ascontiguousarray(s, dtype=np.uint64)
    k = np.ascontiguousarray(k, dtype=np.uint64)
    assert s.shape == k.shape and s.shape[1] == 4
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_dot_mod_r(_ptr(s), _ptr(k), len(s), _ptr(out))
    return words_to_int(out)


def g1_add(a, b):
    x = np.array([a], dtype=G1_AFFINE_DTYPE)
    y = np.array([b], dtype=G1_AFFINE_DTYPE)
    out = np.zeros(1, dtype=G1_AFFINE_DTYPE)
    lib().orc_g1_add(_ptr(out), _ptr(x), _ptr(y))
    return out[0]


def g1_scalar_mul(a, k_int):
    x = np.array([a], dtype=G1_AFFINE_DTYPE)
    k = int_to_words(k_int, 4)
    out = np.zeros(1, dtype=G1_AFFINE_DTYPE)
    lib().orc_g1_scalar_mul(_ptr(out), _ptr(x), _ptr(k))
    return out[0]


def projective_to_affine(p):
    x = np.array([p], dtype=G1_PROJ_DTYPE)
    out = np.zeros(1, dtype=G1_AFFINE_DTYPE)
    lib().orc_g1_projective_to_affine(_ptr(out), _ptr(x))
    return out[0]


def affine_from_json(d):
    a = np.zeros(1, dtype=G1_AFFINE_DTYPE)[0]
    if d["infinity"]:
        a["infinity"] = 1
    else:
        a["x"] = np.array(d["x"], dtype=np.uint64)
        a["y"] = np.array(d["y"], dtype=np.uint64)
    return a


def affine_equal(a, b):
    if bool(a["infinity"]) or bool(b["infinity"]):
        return bool(a["infinity"]) == bool(b["infinity"])
    return np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
