"""Batched and served execution of mixed NTT / MSM jobs on one GPU.

`GpuServer` is a continuous-batching request loop:
  - the MSM base-point tables stay resident in HBM (MsmContext objects);
  - every worker thread owns a HIP stream and an MSM workspace, and
    repeatedly takes the waiting jobs, in arrival order, that fit the free HBM
    budget and an estimated amount of work per batch (`max_batch_s`, so that
    results stream out instead of finishing together), and runs them as ONE
    batch (the NTTs on the worker's persistent plans + one MSM launch sequence);
    requests that
    arrive while other batches are on the GPU are admitted as soon as a worker
    and memory free up, and their work overlaps on the device;
  - the plain form is the same loop with one worker and one job per batch.

No reference counterpart: the reference serialises the device and overlaps
only PCIe transfers (zprize/ntt/host/driver.cpp:148-231,
zprize/msm_pippenger/host/driver/driver.cpp:495-600).
"""
import ctypes
import threading
import time
from collections import deque

import numpy as np

from . import autotune
from ._lib import call, lib
from .msm import G1_PROJ_DTYPE, MsmWorkspace
from .ntt import NttGpuDriver

NTT, MSM = "ntt", "msm"


class Job:
    """One request.  kind NTT: data (uint64 n), direction; kind MSM: table, scalars (n, 4)."""

    __slots__ = ("id", "kind", "log_n", "direction", "data", "table", "scalars", "window", "t_arrival",
                 "t_start", "t_done", "result", "error", "_event")

    def __init__(self, kind, **kw):
        self.kind = kind
        self.id = kw.get("id", 0)
        self.log_n = kw.get("log_n", 0)
        self.direction = kw.get("direction", 0)
        self.data = kw.get("data")
        self.table = kw.get("table", 0)
        self.scalars = kw.get("scalars")
        self.window = kw.get("window", 0)
        self.t_arrival = kw.get("t_arrival", 0.0)
        self.t_start = self.t_done = None
        self.result = None
        self.error = None
        self._event = threading.Event()

    def hbm_bytes(self):
        n = 1 << self.log_n
        if self.kind == NTT:
            # the vector, plus the 4-step scratch of the worker's plan for this
            # size (kept for the worker's life, counted here per job so that
            # the budget stays an upper bound)
            return 16 * n
        # scalars; per (point, window) a digit, a pass-A entry and a sorted
        # entry (16 B); bucket sums, work-unit partials and window groups
        # (192 B points, generous); window c = min(20, max(8, log n - 6))
        c = min(20, max(8, self.log_n - 6))
        w = -(-253 // c)
        return 32 * n + 16 * w * n + 4 * 192 * w * (1 << (c - 1)) + (1 << 26)

    def est_s(self):
        """Rough time of this job inside a batch: host<->device copies of its
        inputs/outputs (pageable, ~25 GB/s) plus device work (MSM ~0.17 ns per
        bucket addition, N per window; NTT ~15 ps per element)."""
        n = 1 << self.log_n
        if self.kind == NTT:
            return 16 * n / 25e9 + 15e-12 * n
        c = min(20, max(8, self.log_n - 6))
        w = -(-254 // c)
        return 32 * n / 25e9 + 0.17e-9 * n * w + 2e-4

    def wait(self, timeout=None):
        self._event.wait(timeout)
        return self


class _DevicePool:
    """Reusable device allocations keyed by size (one pool per worker)."""

    def __init__(self):
        self.free = {}

    def get(self, nbytes):
        lst = self.free.get(nbytes)
        if lst:
            return lst.pop()
        p = ctypes.c_void_p()
        call("zp_device_malloc", ctypes.byref(p), max(nbytes, 1))
        return p.value

    def put(self, nbytes, ptr):
        self.free.setdefault(nbytes, []).append(ptr)

    def release(self):
        for lst in self.free.values():
            for p in lst:
                lib().zp_device_free(ctypes.c_void_p(p))
        self.free.clear()


def admit(queue, in_use, budget, max_jobs, max_batch_s):
    """Pop the next batch from the head of `queue` (a deque of Jobs, arrival
    order): jobs are taken while the batch's HBM need fits `budget - in_use`,
    its estimated work (Job.est_s) stays within `max_batch_s` and it holds at
    most `max_jobs` jobs; the first job is taken whenever it fits the budget,
    whatever its estimated work.  Returns (batch, bytes needed); an empty batch
    when the head job does not fit the free budget."""
    batch, need, work = [], 0.0, 0.0
    while queue and len(batch) < max_jobs:
        b = queue[0].hbm_bytes()
        t = queue[0].est_s()
        if in_use + need + b > budget:
            break
        if batch and work + t > max_batch_s:
            break
        batch.append(queue.popleft())
        need += b
        work += t
    return batch, need


class GpuServer:
    def __init__(self, tables, hbm_budget=200e9, workers=2, max_batch_jobs=256, batching=True, max_batch_s=0.1):
        self.tables = tables
        self.budget = float(hbm_budget)
        self.max_batch_s = float(max_batch_s)
        self.in_use = 0.0
        self.batching = batching
        self.max_batch = max_batch_jobs if batching else 1
        self.queue = deque()
        self.cv = threading.Condition()
        self.stop = False
        self.failed = None  # set when a worker could not start; the server then refuses work
        self.batches = 0
        self.windows = autotune.load()  # size class -> tuned MSM window (empty = library default)
        self.ntt_tiles_applied = autotune.apply_ntt_tiles()  # NTT tiles per size class (process-wide)
        self.threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(workers)]
        for t in self.threads:
            t.start()

    # ---------------------------------------------------------------- API
    def submit(self, job):
        if job.hbm_bytes() > self.budget:
            raise ValueError(f"job {job.id} needs {job.hbm_bytes()} bytes > HBM budget {self.budget}")
        with self.cv:
            if self.failed is not None:
                raise RuntimeError(f"server unavailable: a worker failed to start ({self.failed})")
            self.queue.append(job)
            self.cv.notify()
        return job

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()

    # ------------------------------------------------------------- worker
    def _take(self):
        """Admit waiting jobs (see `admit`); waits while nothing fits."""
        with self.cv:
            while True:
                while not self.queue and not self.stop:
                    self.cv.wait()
                if not self.queue:
                    return None
                batch, need = admit(self.queue, self.in_use, self.budget, self.max_batch, self.max_batch_s)
                if batch:
                    self.in_use += need
                    return batch, need
                self.cv.wait(0.01)  # the head job does not fit the free budget yet

    def _release(self, need):
        with self.cv:
            self.in_use -= need
            self.cv.notify_all()

    def _fail(self, err):
        """A worker could not start: fail every waiting job (so no caller
        waits forever) and refuse new ones."""
        with self.cv:
            if self.failed is None:
                self.failed = err
            pending = list(self.queue)
            self.queue.clear()
            self.cv.notify_all()
        for j in pending:
            j.error = err
            j.t_done = time.perf_counter()
            j._event.set()

    def _worker(self):
        stream = ws = xfer = None
        try:
            s = ctypes.c_void_p()
            call("zp_stream_create", ctypes.byref(s))
            stream = s.value
            ws = MsmWorkspace()
            xfer = _Transfer()
        except Exception as e:
            if stream is not None:
                lib().zp_stream_destroy(ctypes.c_void_p(stream))
            if ws is not None:
                ws.close()
            self._fail(e)
            return
        pool = _DevicePool()
        plans = {}  # log n -> NTT plan (twiddle tables + 4-step scratch), kept for the worker's life
        try:
            while True:
                taken = self._take()
                if taken is None:
                    return
                batch, need = taken
                t0 = time.perf_counter()
                for j in batch:
                    j.t_start = t0
                try:
                    self._run_batch(batch, stream, ws, pool, plans, xfer)
                except Exception as e:  # surfaced on the job
                    for j in batch:
                        j.error = e
                t1 = time.perf_counter()
                self._release(need)
                with self.cv:
                    self.batches += 1
                for j in batch:
                    j.t_done = t1
                    j._event.set()
        finally:
            ws.close()
            xfer.close()
            for p in plans.values():
                p.close()
            pool.release()
            lib().zp_stream_destroy(ctypes.c_void_p(stream))

    def _run_batch(self, batch, stream, ws, pool, plans, xfer):
        """One batch on the worker's stream: inputs to HBM through the
        worker's pinned staging, the NTTs on the worker's persistent plans
        (no allocation per batch), the MSMs as ONE launch sequence, NTT
        results back through the staging.  The stream is synchronized and the
        pool buffers returned whatever fails."""
        ntt_jobs = [j for j in batch if j.kind == NTT]
        msm_jobs = [j for j in batch if j.kind == MSM]
        held = []
        sp = ctypes.c_void_p(stream)
        try:
            ntt_ptrs, msm_ptrs = [], []
            for j in ntt_jobs:
                nb = 8 << j.log_n
                p = pool.get(nb)
                held.append((nb, p))
                ntt_ptrs.append(p)
                data = np.ascontiguousarray(j.data, dtype=np.uint64)
                assert data.nbytes == nb, "NTT job: data length is not 2^log_n"
                xfer.h2d(p, data)
            for j in msm_jobs:
                nb = 32 * max(len(j.scalars), 1)
                p = pool.get(nb)
                held.append((nb, p))
                msm_ptrs.append(p)
                xfer.h2d(p, np.ascontiguousarray(j.scalars, dtype=np.uint64))
            for j, p in zip(ntt_jobs, ntt_ptrs):
                plan = plans.get(j.log_n)
                if plan is None:
                    plan = plans[j.log_n] = NttGpuDriver(j.log_n, num_buffers=1)
                plan.inplace_device(p, j.direction, stream)
            if msm_jobs:
                specs = [(self.tables[j.table], len(j.scalars), p,
                          j.window or self.windows.get(max(len(j.scalars) - 1, 1).bit_length(), 0))
                         for j, p in zip(msm_jobs, msm_ptrs)]
                out = ws.run(specs, stream=stream)
                for j, o in zip(msm_jobs, out):
                    j.result = o
            for j, p in zip(ntt_jobs, ntt_ptrs):
                res = np.empty(1 << j.log_n, dtype=np.uint64)
                xfer.d2h(res, p, after_stream=stream)
                j.result = res
        finally:
            lib().zp_stream_synchronize(sp)
            for nb, p in held:
                pool.put(nb, p)


class _Transfer:
    """zp_transfer: pinned multi-lane staging for large host arrays."""

    def __init__(self):
        h = ctypes.c_void_p()
        call("zp_transfer_create", ctypes.byref(h))
        self.h = h

    def h2d(self, d_ptr, arr):
        call("zp_transfer_h2d", self.h, ctypes.c_void_p(d_ptr), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes)

    def d2h(self, arr, d_ptr, after_stream=None):
        call("zp_transfer_d2h", self.h, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(d_ptr), arr.nbytes,
             ctypes.c_void_p(after_stream or 0))

    def close(self):
        if self.h:
            lib().zp_transfer_destroy(self.h)
            self.h = None


def run_stream(server, jobs, arrivals):
    """Submit jobs at the given arrival offsets (seconds from now), wait for all.
    Returns (makespan_s, latencies_s)."""
    t0 = time.perf_counter()
    for j, a in zip(jobs, arrivals):
        now = time.perf_counter() - t0
        if a > now:
            time.sleep(a - now)
        j.t_arrival = time.perf_counter()
        server.submit(j)
    for j in jobs:
        j.wait()
    t1 = max(j.t_done for j in jobs)
    for j in jobs:
        if j.error is not None:
            raise j.error
    return t1 - t0, [j.t_done - j.t_arrival for j in jobs]


def percentiles(values, ps=(50, 99)):
    v = np.sort(np.asarray(values))
    return {f"p{p}": float(np.percentile(v, p)) for p in ps}


__all__ = ["GpuServer", "Job", "NTT", "MSM", "admit", "run_stream", "percentiles", "G1_PROJ_DTYPE"]
