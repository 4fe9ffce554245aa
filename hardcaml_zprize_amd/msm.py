"""BLS12-377 G1 MSM host driver over the C ABI.

Mirrors the reference's two-call interface (zprize/msm_pippenger/host/driver/driver.cpp:603-636,
used by test_fpga_harness/src/lib.rs:40-85): load the base points once
(`MsmContext(points)` == msm_init), then run `mult(scalars)` (== msm_mult)
over fresh scalars, one output point per batch.
"""
import ctypes

import numpy as np

from ._lib import call, lib

# rust_types.h:63-83 layouts.
G1_AFFINE_DTYPE = np.dtype([("x", "<u8", (6,)), ("y", "<u8", (6,)), ("infinity", "u1"), ("_pad", "u1", (7,))])
G1_PROJ_DTYPE = np.dtype([("x", "<u8", (6,)), ("y", "<u8", (6,)), ("z", "<u8", (6,))])
assert G1_AFFINE_DTYPE.itemsize == 104 and G1_PROJ_DTYPE.itemsize == 144


def _vp(x):
    return ctypes.c_void_p(int(x) if x else 0)


class MsmContext:
    """msm_init: converts and keeps the base points resident in HBM.

    points: a host array of G1_AFFINE_DTYPE, or (device_ptr, n) for points
    already in HBM.  options: e.g. "window=reference", "window=16".
    """

    def __init__(self, points, options=""):
        h = ctypes.c_void_p()
        opt = options.encode()
        if isinstance(points, tuple):
            d_ptr, n = points
            call("zp_msm_init_device", opt, len(opt), _vp(d_ptr), int(n), ctypes.byref(h))
            self.n = int(n)
        else:
            pts = np.ascontiguousarray(points, dtype=G1_AFFINE_DTYPE)
            call("zp_msm_init", opt, len(opt), pts.ctypes.data_as(ctypes.c_void_p), len(pts), ctypes.byref(h))
            self.n = len(pts)
        self.ctx = h

    def close(self):
        if self.ctx:
            lib().msm_free(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def precompute_width(self):
        """Width of the window-shifted table (init option precompute=), 0 if none."""
        return int(lib().zp_msm_precompute_width(self.ctx))

    def plan(self, n=None):
        """(shifted, window_bits, num_windows) of an MSM over the first n points."""
        sh, wb, nw = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call("zp_debug_msm_plan", self.ctx, self.n if n is None else int(n), ctypes.byref(sh), ctypes.byref(wb),
             ctypes.byref(nw))
        return bool(sh.value), wb.value, nw.value

    def num_nonconvertible(self):
        return int(lib().zp_msm_num_nonconvertible(self.ctx))

    def mult(self, scalars, num_batches=None):
        """scalars: uint64 array (num_batches * n, 4); returns G1_PROJ_DTYPE[num_batches]."""
        sc = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
        if num_batches is None:
            num_batches = len(sc) // max(self.n, 1) if self.n else 1
        assert len(sc) == num_batches * self.n
        out = np.zeros(num_batches, dtype=G1_PROJ_DTYPE)
        call("zp_msm_mult", self.ctx, out.ctypes.data_as(ctypes.c_void_p), num_batches,
             sc.ctypes.data_as(ctypes.c_void_p))
        return out

    def mult_abi(self, scalars, num_batches):
        """The void msm_mult entry point itself (errors only print)."""
        sc = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
        out = np.zeros(num_batches, dtype=G1_PROJ_DTYPE)
        lib().msm_mult(self.ctx, out.ctypes.data_as(ctypes.c_void_p), num_batches, sc.ctypes.data_as(ctypes.c_void_p))
        return out

    def mult_device(self, d_scalars, num_batches=1, stream=None):
        out = np.zeros(num_batches, dtype=G1_PROJ_DTYPE)
        call("zp_msm_mult_device", self.ctx, out.ctypes.data_as(ctypes.c_void_p), num_batches, _vp(d_scalars),
             _vp(stream))
        return out

    def mult_prefix_device(self, n, d_scalars, stream=None):
        out = np.zeros(1, dtype=G1_PROJ_DTYPE)
        call("zp_msm_mult_prefix_device", self.ctx, int(n), out.ctypes.data_as(ctypes.c_void_p), _vp(d_scalars),
             _vp(stream))
        return out[0]

    def precomputed(self, i0, count):
        out = np.zeros((count, 18), dtype=np.uint64)
        call("zp_debug_msm_precomputed", self.ctx, i0, count, out.ctypes.data_as(ctypes.c_void_p))
        return out


class MsmJob(ctypes.Structure):
    """zp_msm_job: one MSM over points [0, n) of a resident table."""
    _fields_ = [("ctx", ctypes.c_void_p), ("n", ctypes.c_size_t), ("d_scalars", ctypes.c_void_p),
                ("out", ctypes.c_void_p), ("window", ctypes.c_int)]


class MsmWorkspace:
    """Device working memory for one batch at a time (zp_msm_workspace)."""

    def __init__(self):
        h = ctypes.c_void_p()
        call("zp_msm_workspace_create", ctypes.byref(h))
        self.h = h

    def close(self):
        if self.h:
            lib().zp_msm_workspace_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, jobs, options="", stream=None):
        """jobs: list of (MsmContext, n, d_scalars_ptr[, window]); returns G1_PROJ_DTYPE[len(jobs)]."""
        out = np.zeros(len(jobs), dtype=G1_PROJ_DTYPE)
        arr = (MsmJob * len(jobs))()
        base = out.ctypes.data
        for i, job in enumerate(jobs):
            ctx, n, dsc = job[:3]
            window = int(job[3]) if len(job) > 3 and job[3] else 0
            arr[i] = MsmJob(ctx.ctx, int(n), int(dsc) if dsc else 0, base + i * G1_PROJ_DTYPE.itemsize, window)
        opt = options.encode()
        call("zp_msm_batch_device", self.h, arr, len(jobs), opt, len(opt), _vp(stream))
        return out


class DeviceBuffer:
    """A plain HBM allocation through the library (no torch needed)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        call("zp_device_malloc", ctypes.byref(p), max(int(nbytes), 1))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        call("zp_memcpy_h2d", _vp(self.ptr), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes)
        return self

    def download(self, arr):
        call("zp_memcpy_d2h", arr.ctypes.data_as(ctypes.c_void_p), _vp(self.ptr), arr.nbytes)
        return arr

    def free(self):
        if self.ptr:
            call("zp_device_free", _vp(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_points_device(d_k, d_out, n, stream=None):
    """out[i] = k_i * G on the GPU (wire-form affine points)."""
    call("zp_g1_mul_generator_device", _vp(d_k), _vp(d_out), int(n), _vp(stream))


def generate_points(ks):
    """Host convenience: k_i * G for a uint64 (n, 4) array, via the GPU."""
    ks = np.ascontiguousarray(ks, dtype=np.uint64).reshape(-1, 4)
    n = len(ks)
    dk = DeviceBuffer(32 * n).upload(ks)
    dp = DeviceBuffer(104 * n)
    generate_points_device(dk.ptr, dp.ptr, n)
    out = np.zeros(n, dtype=G1_AFFINE_DTYPE)
    dp.download(out)
    dk.free()
    dp.free()
    return out


def debug_fq_mul(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 6)
    b = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, 6)
    out = np.zeros_like(a)
    call("zp_debug_fq_mul", a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
         out.ctypes.data_as(ctypes.c_void_p), len(a))
    return out


def debug_digits(scalars, options=""):
    sc = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    opt = options.encode()
    W = ctypes.c_int()
    call("zp_debug_msm_digits", opt, len(opt), None, len(sc), None, ctypes.byref(W))
    out = np.zeros((len(sc), W.value), dtype=np.int32)
    call("zp_debug_msm_digits", opt, len(opt), sc.ctypes.data_as(ctypes.c_void_p), len(sc),
         out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(W))
    return out
