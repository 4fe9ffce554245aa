"""Loader for the native library hardcaml_zprize_amd/libzpgpu.so (C ABI of include/zpgpu.h).

There is no Python or CPU fallback: if the library is missing or no GPU is
present, the calls fail loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzpgpu.so")

ZP_OK = 0
STATUS_NAMES = {0: "ZP_OK", 1: "ZP_ERR_INVALID_ARG", 2: "ZP_ERR_HIP", 3: "ZP_ERR_OUT_OF_MEMORY",
                4: "ZP_ERR_NO_BUFFER", 5: "ZP_ERR_STATE", 6: "ZP_ERR_TOO_LARGE", 7: "ZP_ERR_NO_DEVICE"}

# Every symbol include/zpgpu.h declares, with (restype, argtypes).
_vp, _u64, _u32, _i32, _sz = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint, ctypes.c_int, ctypes.c_size_t
SIGNATURES = {
    "zp_last_error": (ctypes.c_char_p, []),
    "zp_version": (ctypes.c_char_p, []),
    "zp_ntt_create": (_i32, [_u32, _u32, ctypes.POINTER(_vp)]),
    "zp_ntt_create_reverse": (_i32, [_u32, _u32, ctypes.POINTER(_vp)]),
    "zp_ntt_destroy": (_i32, [_vp]),
    "zp_ntt_size": (_u64, [_vp]),
    "zp_ntt_max_buffers_in_flight": (_u32, [_vp]),
    "zp_ntt_inplace_device": (_i32, [_vp, _vp, _i32, _vp]),
    "zp_ntt_time_passes": (_i32, [_vp, _vp, _i32, _i32, _vp, ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_float)]),
    "zp_ntt_simple_evaluate": (_i32, [_vp, _vp, _vp, _u64, _i32]),
    "zp_ntt_request_buffer": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "zp_ntt_buffer_input": (_vp, [_vp, _i32]),
    "zp_ntt_buffer_output": (_vp, [_vp, _i32]),
    "zp_ntt_enqueue": (_i32, [_vp, _i32, _i32]),
    "zp_ntt_wait": (_i32, [_vp, _i32]),
    "zp_ntt_free_buffer": (_i32, [_vp, _i32]),
    "zp_ntt_batch_device": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "zp_ntt_batch_create": (_i32, [_vp, _vp, _vp, _sz, ctypes.POINTER(_vp)]),
    "zp_ntt_batch_launch": (_i32, [_vp, _i32, _vp]),
    "zp_ntt_batch_destroy": (_i32, [_vp]),
    "zp_ntt_set_tiles": (_i32, [_u32, _i32, _i32, _i32, _i32]),
    "zp_ntt_get_tiles": (_i32, [_u32, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "msm_init": (_vp, [ctypes.c_char_p, ctypes.c_ssize_t, _vp, ctypes.c_ssize_t]),
    "msm_mult": (None, [_vp, _vp, _u64, _vp]),
    "msm_free": (None, [_vp]),
    "zp_msm_init": (_i32, [ctypes.c_char_p, _sz, _vp, _sz, ctypes.POINTER(_vp)]),
    "zp_msm_init_device": (_i32, [ctypes.c_char_p, _sz, _vp, _sz, ctypes.POINTER(_vp)]),
    "zp_msm_mult": (_i32, [_vp, _vp, _u64, _vp]),
    "zp_fr_to_montgomery_device": (_i32, [_vp, _u64, _vp]),
    "zp_msm_mult_device": (_i32, [_vp, _vp, _u64, _vp, _vp]),
    "zp_msm_mult_prefix_device": (_i32, [_vp, _sz, _vp, _vp, _vp]),
    "zp_msm_mult_prefix_noalloc": (_i32, [_vp, _sz, _vp, _vp, _vp]),
    "zp_msm_num_points": (_sz, [_vp]),
    "zp_msm_precompute_width": (_i32, [_vp]),
    "zp_debug_msm_plan": (_i32, [_vp, _sz, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "zp_msm_num_nonconvertible": (_sz, [_vp]),
    "zp_msm_workspace_create": (_i32, [ctypes.POINTER(_vp)]),
    "zp_msm_workspace_destroy": (_i32, [_vp]),
    "zp_msm_batch_device": (_i32, [_vp, _vp, _sz, ctypes.c_char_p, _sz, _vp]),
    "zp_mixed_batch_create": (_i32, [_vp, _vp, _vp, _sz, _vp, _sz, ctypes.c_char_p, _sz, ctypes.POINTER(_vp)]),
    "zp_mixed_batch_launch": (_i32, [_vp, _vp]),
    "zp_mixed_batch_wait": (_i32, [_vp, _vp]),
    "zp_mixed_batch_destroy": (_i32, [_vp]),
    "zp_g1_mul_generator_device": (_i32, [_vp, _vp, _sz, _vp]),
    "zp_debug_fq_mul": (_i32, [_vp, _vp, _vp, _sz]),
    "zp_debug_msm_digits": (_i32, [ctypes.c_char_p, _sz, _vp, _sz, _vp, ctypes.POINTER(_i32)]),
    "zp_debug_msm_precomputed": (_i32, [_vp, _sz, _sz, _vp]),
    "zp_probe_valu_rate": (_i32, [_i32, ctypes.POINTER(ctypes.c_double)]),
    "zp_device_malloc": (_i32, [ctypes.POINTER(_vp), _sz]),
    "zp_device_free": (_i32, [_vp]),
    "zp_memcpy_h2d": (_i32, [_vp, _vp, _sz]),
    "zp_memcpy_d2h": (_i32, [_vp, _vp, _sz]),
    "zp_device_synchronize": (_i32, []),
    "zp_stream_create": (_i32, [ctypes.POINTER(_vp)]),
    "zp_stream_destroy": (_i32, [_vp]),
    "zp_stream_synchronize": (_i32, [_vp]),
    "zp_memcpy_h2d_async": (_i32, [_vp, _vp, _sz, _vp]),
    "zp_memcpy_d2h_async": (_i32, [_vp, _vp, _sz, _vp]),
    "zp_host_alloc": (_i32, [ctypes.POINTER(_vp), _sz]),
    "zp_host_free": (_i32, [_vp]),
    "zp_transfer_create": (_i32, [ctypes.POINTER(_vp)]),
    "zp_transfer_destroy": (_i32, [_vp]),
    "zp_transfer_h2d": (_i32, [_vp, _vp, _vp, _sz]),
    "zp_transfer_d2h": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "zp_device_memory_info": (_i32, [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
}

_lib = None


class ZpError(RuntimeError):
    def __init__(self, fn, status, text):
        super().__init__(f"{fn} -> {STATUS_NAMES.get(status, status)}: {text}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library {LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue  # reported by missing_symbols(); calling it raises AttributeError
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def missing_symbols():
    """Names from include/zpgpu.h that the built library does not export."""
    L = lib()
    return [n for n in SIGNATURES if getattr(L, n, None) is None]


def last_error():
    t = lib().zp_last_error()
    return t.decode() if t else ""


def check(fn_name, status):
    if status != ZP_OK:
        raise ZpError(fn_name, status, last_error())
    return status


def call(fn_name, *args):
    return check(fn_name, getattr(lib(), fn_name)(*args))
