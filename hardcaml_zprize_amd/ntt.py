"""Goldilocks NTT host driver over the C ABI.

`NttGpuDriver` keeps the method names, argument meaning and error behaviour of
the reference's `NttFpgaDriver` (zprize/ntt/host/driver.h:58-225):
request_buffer returns None when every buffer is in flight, free_buffer of a
free buffer raises, simple_evaluate zero-pads short input and raises when the
input is longer than the transform.
"""
import ctypes

import numpy as np

from ._lib import ZpError, call, lib

FORWARD, INVERSE = 0, 1


class UserBuffer:
    """One in-flight buffer (driver.h:62-93): pinned host input/output vectors."""

    def __init__(self, driver, buffer_id):
        self.driver = driver
        self.id = buffer_id
        n = driver.input_vector_size()
        self._in = _host_view(lib().zp_ntt_buffer_input(driver.plan, buffer_id), n)
        self._out = _host_view(lib().zp_ntt_buffer_output(driver.plan, buffer_id), n)

    def input_data(self):
        return self._in

    def output_data(self):
        return self._out


def _host_view(addr, n):
    if not addr:
        raise RuntimeError("buffer pointer is NULL")
    buf = (ctypes.c_uint64 * n).from_address(addr)
    return np.frombuffer(buf, dtype=np.uint64, count=n)


class NttGpuDriver:
    """NttFpgaDriver(create_ntt_2_<log_n>) + load_xclbin in one constructor.
    `NttGpuDriver.reverse(log_row_size)` is NttFpgaDriverArg::create_reverse:
    the REVERSE debug core on 4^log_row_size elements."""

    def __init__(self, log_n, num_buffers=8, _reverse=False):
        h = ctypes.c_void_p()
        if _reverse:
            call("zp_ntt_create_reverse", log_n // 2, num_buffers, ctypes.byref(h))
        else:
            call("zp_ntt_create", log_n, num_buffers, ctypes.byref(h))
        self.plan = h
        self.log_n = log_n

    @classmethod
    def reverse(cls, log_row_size, num_buffers=8):
        return cls(2 * log_row_size, num_buffers, _reverse=True)

    def close(self):
        if self.plan:
            lib().zp_ntt_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def input_vector_size(self):
        return int(lib().zp_ntt_size(self.plan))

    def max_num_buffers_in_flight(self):
        return int(lib().zp_ntt_max_buffers_in_flight(self.plan))

    def request_buffer(self):
        bid = ctypes.c_int(-1)
        st = lib().zp_ntt_request_buffer(self.plan, ctypes.byref(bid))
        if st == 4:  # ZP_ERR_NO_BUFFER: the reference returns NULL
            return None
        if st != 0:
            call("zp_ntt_request_buffer", st)
        return UserBuffer(self, bid.value)

    def transfer_input_vector_to_user_buffer(self, buf, data):
        data = np.asarray(data, dtype=np.uint64)
        buf.input_data()[:] = data

    def enqueue_evaluation_async(self, buf, direction=FORWARD):
        call("zp_ntt_enqueue", self.plan, buf.id, direction)

    def wait_for_result(self, buf):
        call("zp_ntt_wait", self.plan, buf.id)

    def transfer_user_buffer_to_output_vector(self, buf, out=None):
        if out is None:
            return buf.output_data().copy()
        out[:] = buf.output_data()
        return out

    def free_buffer(self, buf):
        call("zp_ntt_free_buffer", self.plan, buf.id)

    def simple_evaluate(self, data, direction=FORWARD):
        data = np.ascontiguousarray(data, dtype=np.uint64)
        out = np.empty(self.input_vector_size(), dtype=np.uint64)
        call("zp_ntt_simple_evaluate", self.plan, out.ctypes.data, data.ctypes.data, len(data), direction)
        return out

    def inplace_device(self, d_ptr, direction=FORWARD, stream=None):
        """The hot path: in-place transform of n u64 already in HBM."""
        call("zp_ntt_inplace_device", self.plan, ctypes.c_void_p(d_ptr), direction, ctypes.c_void_p(stream or 0))


    def time_passes(self, d_ptr, direction=FORWARD, reps=10, stream=None):
        """Mean milliseconds of the transform's two kernel launches (pass 1,
        pass 2), timed with HIP events on `stream` (zp_ntt_time_passes)."""
        m1, m2 = ctypes.c_float(), ctypes.c_float()
        call("zp_ntt_time_passes", self.plan, ctypes.c_void_p(d_ptr), direction, int(reps),
             ctypes.c_void_p(stream or 0), ctypes.byref(m1), ctypes.byref(m2))
        return m1.value, m2.value


def ntt_batch_device(ptrs, log_ns, dirs, stream=None):
    """Many in-place transforms of mixed sizes/directions, one launch sequence."""
    n = len(ptrs)
    P = (ctypes.c_void_p * n)(*ptrs)
    L = (ctypes.c_uint * n)(*log_ns)
    D = (ctypes.c_int * n)(*dirs)
    call("zp_ntt_batch_device", P, L, D, n, ctypes.c_void_p(stream or 0))


TILE_MIN_LOG, TILE_MAX_LOG = 10, 23


def set_tiles(log_n, batched, log_n1=-1, log_t1=-1, log_t2=-1):
    """Tile choice of one size class (log_n in [10, 23]) for single transforms
    (batched=False) or batch groups (batched=True): column length 2^log_n1
    (2^12: pass 1 on the register-resident column kernel) and log2 sequences
    per workgroup of each pass; -1 = the built-in policy.  Raises ZpError on
    an invalid choice (sub-DFT lengths outside 2^1..2^12, a tile beyond 2^13
    elements)."""
    call("zp_ntt_set_tiles", int(log_n), 1 if batched else 0, int(log_n1), int(log_t1), int(log_t2))


def get_tiles(log_n, batched):
    """Effective (log_n1, log_t1, log_t2) of a size class."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("zp_ntt_get_tiles", int(log_n), 1 if batched else 0, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def tile_candidates(log_n):
    """Every valid (log_n1, log_t1, log_t2) of a size class."""
    out = []
    for n1 in range(max(1, log_n - 12), min(12, log_n - 1) + 1):
        n2 = log_n - n1
        for t1 in range(0, min(13 - n1, n2) + 1):
            for t2 in range(0, min(13 - n2, n1) + 1):
                out.append((n1, t1, t2))
    return out


class NttBatch:
    """Many in-place transforms of mixed sizes/directions, grouped into one
    launch per (size, direction, pass); optionally replayed from a HIP graph."""

    def __init__(self, ptrs, log_ns, dirs):
        n = len(ptrs)
        P = (ctypes.c_void_p * n)(*ptrs)
        L = (ctypes.c_uint * n)(*log_ns)
        D = (ctypes.c_int * n)(*dirs)
        h = ctypes.c_void_p()
        call("zp_ntt_batch_create", P, L, D, n, ctypes.byref(h))
        self.h = h

    def launch(self, use_graph=True, stream=None):
        call("zp_ntt_batch_launch", self.h, 1 if use_graph else 0, ctypes.c_void_p(stream or 0))

    def close(self):
        if self.h:
            lib().zp_ntt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["NttBatch", "NttGpuDriver", "UserBuffer", "ntt_batch_device", "set_tiles", "get_tiles", "tile_candidates",
           "FORWARD", "INVERSE", "ZpError"]
