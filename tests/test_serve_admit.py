"""Host logic of the served form: batch admission (hardcaml_zprize_amd/serve.py
`admit`).  No GPU: only the queue policy is exercised -- arrival order, the
HBM budget, the per-batch work estimate and the job-count cap."""
from collections import deque

import pytest

from hardcaml_zprize_amd import serve


def msm(i, log_n):
    return serve.Job(serve.MSM, id=i, log_n=log_n)


def ntt(i, log_n):
    return serve.Job(serve.NTT, id=i, log_n=log_n)


def ids(batch):
    return [j.id for j in batch]


def test_estimates_grow_with_size():
    assert msm(0, 20).est_s() > msm(0, 14).est_s() > 0
    assert ntt(0, 24).est_s() > ntt(0, 12).est_s() > 0
    # a 2^24 MSM: 2^24 points x 15 windows of bucket additions plus 512 MB of scalars
    assert 0.03 < msm(0, 24).est_s() < 0.1


def test_takes_every_job_when_unbounded():
    q = deque([msm(0, 14), ntt(1, 12), msm(2, 16)])
    batch, need = serve.admit(q, 0.0, 200e9, 256, float("inf"))
    assert ids(batch) == [0, 1, 2] and not q
    assert need == sum(j.hbm_bytes() for j in batch)


def test_work_cap_keeps_arrival_order():
    jobs = [msm(i, 20) for i in range(6)]
    per = jobs[0].est_s()
    q = deque(jobs)
    batch, _ = serve.admit(q, 0.0, 200e9, 256, 2.5 * per)
    assert ids(batch) == [0, 1]  # a third job would exceed 2.5 jobs' worth
    batch, _ = serve.admit(q, 0.0, 200e9, 256, 2.5 * per)
    assert ids(batch) == [2, 3]
    assert ids(q) == [4, 5]


def test_first_job_ignores_work_cap():
    q = deque([msm(0, 24), ntt(1, 12)])
    batch, _ = serve.admit(q, 0.0, 200e9, 256, 1e-6)
    assert ids(batch) == [0]  # a batch always makes progress
    batch, _ = serve.admit(q, 0.0, 200e9, 256, 1e-6)
    assert ids(batch) == [1]


def test_budget_blocks_head_and_caps_batch():
    big, small = msm(0, 24), ntt(1, 12)
    q = deque([big, small])
    # head does not fit the free memory: nothing is taken (the worker waits)
    batch, need = serve.admit(q, 200e9 - big.hbm_bytes() / 2, 200e9, 256, float("inf"))
    assert batch == [] and need == 0.0 and len(q) == 2
    # room for the head only
    batch, need = serve.admit(q, 0.0, big.hbm_bytes() + small.hbm_bytes() / 2, 256, float("inf"))
    assert ids(batch) == [0] and need == big.hbm_bytes()


@pytest.mark.parametrize("cap", [1, 3])
def test_job_count_cap(cap):
    q = deque(ntt(i, 10) for i in range(5))
    batch, _ = serve.admit(q, 0.0, 200e9, cap, float("inf"))
    assert ids(batch) == list(range(cap))


def test_worker_start_failure_fails_waiting_jobs(monkeypatch):
    """A worker that cannot create its stream fails the jobs already queued
    (their waiters return with the error instead of blocking forever) and the
    server refuses later submissions.  The failure is forced on the host."""
    import threading

    release = threading.Event()

    def failing_call(name, *args):
        release.wait(10)
        raise RuntimeError(f"{name}: no device")

    monkeypatch.setattr(serve, "call", failing_call)
    srv = serve.GpuServer({}, workers=1)
    queued = srv.submit(ntt(0, 10))  # the worker is still starting
    release.set()
    assert queued._event.wait(10), "queued job was never completed"
    assert isinstance(queued.error, RuntimeError)
    srv.threads[0].join(10)
    with pytest.raises(RuntimeError, match="worker failed to start"):
        srv.submit(ntt(1, 10))
    srv.close()
