"""The batched form: many NTTs and MSMs of uneven sizes as ONE HIP graph.

`MixedBatch` plans, allocates and captures the whole batch once
(zp_mixed_batch_create): the NTT groups (one launch per pass per
(log n, direction) group) and the MSM launch sequence (every job's windows
and buckets concatenated) are two parallel branches of one graph, whose
last node copies the MSM results to pinned host memory.  `launch()` replays
the graph (asynchronous), `wait()` returns the MSM results.  Per-job MSM
window widths and the NTT groups' tiles come from the autotuner's table
(`autotune.window_for`, `autotune.apply_ntt_tiles`).

No reference counterpart: the reference runs one transform or one MSM at a
time on its device (zprize/ntt/host/bench_throughput.cpp:61-107,
zprize/msm_pippenger/host/driver/driver.cpp:495-600).
"""
import ctypes

import numpy as np

from . import autotune
from ._lib import call, lib
from .msm import G1_PROJ_DTYPE, MsmJob


def _vp(x):
    return ctypes.c_void_p(int(x) if x else 0)


class MixedBatch:
    """ntt: list of (device_ptr, log_n, direction); msm: list of
    (MsmContext, n, d_scalars_ptr[, window]).  window 0 or absent: the
    autotuner's width for the job's size class (0 = library default) when
    `tuned`, else the library default.  `tuned` also sets the library's NTT
    tiles from the autotuner's table before the graph is captured."""

    def __init__(self, ntt, msm, options="", tuned=True, table=None):
        self.out = np.zeros(max(len(msm), 1), dtype=G1_PROJ_DTYPE)
        windows = autotune.load() if (tuned and table is None) else (table or {})
        self.ntt_tiles_applied = autotune.apply_ntt_tiles() if tuned else 0
        self.windows = []
        jobs = (MsmJob * max(len(msm), 1))()
        base = self.out.ctypes.data
        for i, job in enumerate(msm):
            ctx, n, dsc = job[:3]
            w = int(job[3]) if len(job) > 3 and job[3] else windows.get(max(int(n) - 1, 1).bit_length(), 0)
            self.windows.append(w)
            jobs[i] = MsmJob(ctx.ctx, int(n), int(dsc) if dsc else 0, base + i * G1_PROJ_DTYPE.itemsize, int(w))
        k = len(ntt)
        ptrs = (ctypes.c_void_p * max(k, 1))(*[int(p) for p, _, _ in ntt])
        logs = (ctypes.c_uint * max(k, 1))(*[int(lg) for _, lg, _ in ntt])
        dirs = (ctypes.c_int * max(k, 1))(*[int(d) for _, _, d in ntt])
        opt = options.encode()
        h = ctypes.c_void_p()
        call("zp_mixed_batch_create", ptrs, logs, dirs, k, jobs, len(msm), opt, len(opt), ctypes.byref(h))
        self.h = h
        self.msm_count = len(msm)

    def launch(self, stream=None):
        call("zp_mixed_batch_launch", self.h, _vp(stream))

    def wait(self, stream=None):
        call("zp_mixed_batch_wait", self.h, _vp(stream))
        return self.out[:self.msm_count].copy()

    def run(self, stream=None):
        self.launch(stream)
        return self.wait(stream)

    def close(self):
        if self.h:
            lib().zp_mixed_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
