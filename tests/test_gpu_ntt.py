"""Goldilocks NTT on the GPU through the C ABI, bit-exact against the oracle."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import ntt as zntt
from hardcaml_zprize_amd._lib import call

pytestmark = pytest.mark.gpu
P = oracle.P


def _rand(n, seed):
    return np.random.default_rng(seed).integers(0, P, size=n, dtype=np.uint64)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_kats(ntt_golden):
    _say("ntt kats")
    for name in ("kat_8pt_linear", "kat_8pt_random"):
        kat = ntt_golden[name]
        with zntt.NttGpuDriver(3) as d:
            assert list(d.simple_evaluate(kat["input"])) == kat["output"]
    kat = ntt_golden["kat_16pt_4step"]
    with zntt.NttGpuDriver(4) as d:
        assert list(d.simple_evaluate(kat["input"])) == kat["output"]
    with zntt.NttGpuDriver(12) as d:
        out = d.simple_evaluate([1, 2])  # zero padded to 4096
        assert list(out[:8]) == ntt_golden["kat_4096pt_linear_first8"]["output_first8"]


@pytest.mark.parametrize("log_n", list(range(1, 21)))
def test_random_vs_oracle(log_n):
    _say(f"ntt random log_n={log_n}")
    a = _rand(1 << log_n, 1000 + log_n)
    with zntt.NttGpuDriver(log_n) as d:
        fwd = d.simple_evaluate(a, zntt.FORWARD)
        assert np.array_equal(fwd, oracle.ntt(a))
        inv = d.simple_evaluate(a, zntt.INVERSE)
        assert np.array_equal(inv, oracle.ntt(a, inverse=True))


@pytest.mark.parametrize("log_n", [21, 22, 23, 24])
def test_large_vs_oracle(log_n):
    # the 2^24 workload: a 4096 x 4096 4-step, seed 1, bit-exact vs the CPU reference
    _say(f"ntt large log_n={log_n}")
    a = _rand(1 << log_n, 1 if log_n == 24 else log_n)
    with zntt.NttGpuDriver(log_n) as d:
        fwd = d.simple_evaluate(a, zntt.FORWARD)
        assert np.array_equal(fwd, oracle.ntt(a))
        back = d.simple_evaluate(fwd, zntt.INVERSE)
        assert np.array_equal(back, a)


@pytest.mark.parametrize("log_n,tiles", [(3, None), (10, None), (13, None), (17, None), (17, (8, 2, 1)),
                                         (20, None), (21, (11, 1, 2)), (22, None), (23, None), (24, None)])
def test_noncanonical_inputs_count_as_residues(log_n, tiles):
    """Inputs in [p, 2^64) transform as their residue mod p at every size
    (the single-pass, radix-2 LDS, register-stage LDS and register-resident
    paths alike; the tile choices put pass 1 on the register-stage kernel)."""
    _say(f"ntt noncanonical log_n={log_n} tiles={tiles}")
    if tiles:
        zntt.set_tiles(log_n, False, *tiles)
    n = 1 << log_n
    rng = np.random.default_rng(500 + log_n)
    a = rng.integers(0, P, size=n, dtype=np.uint64)
    big = a.copy()
    sel = rng.random(n) < 0.5
    big[sel] = (a[sel] % np.uint64(0xFFFFFFFF)) + np.uint64(P)  # p + r, r < 2^32 - 1: still a u64, = r mod p
    red = big.copy()
    red[sel] = big[sel] - np.uint64(P)
    try:
        with zntt.NttGpuDriver(log_n) as d:
            assert np.array_equal(d.simple_evaluate(big, zntt.FORWARD), oracle.ntt(red))
            assert np.array_equal(d.simple_evaluate(big, zntt.INVERSE), oracle.ntt(red, inverse=True))
    finally:
        if tiles:
            zntt.set_tiles(log_n, False)


def test_edge_values():
    _say("ntt edge values")
    with zntt.NttGpuDriver(14) as d:
        n = 1 << 14
        for a in (np.zeros(n, dtype=np.uint64), np.full(n, P - 1, dtype=np.uint64), np.ones(n, dtype=np.uint64)):
            assert np.array_equal(d.simple_evaluate(a), oracle.ntt(a))
        ones = d.simple_evaluate(np.ones(n, dtype=np.uint64))
        assert int(ones[0]) == n and not ones[1:].any()


def test_buffer_api_and_errors():
    _say("ntt buffer api")
    with zntt.NttGpuDriver(13, num_buffers=8) as d:
        assert d.max_num_buffers_in_flight() == 8
        bufs = [d.request_buffer() for _ in range(8)]
        assert all(b is not None for b in bufs)
        assert d.request_buffer() is None  # driver.h:146-150: NULL when all in use
        inputs = [_rand(1 << 13, 50 + i) for i in range(8)]
        for b, a, i in zip(bufs, inputs, range(8)):
            d.transfer_input_vector_to_user_buffer(b, a)
            d.enqueue_evaluation_async(b, i % 2)
        for b, a, i in zip(bufs, inputs, range(8)):
            d.wait_for_result(b)
            out = d.transfer_user_buffer_to_output_vector(b)
            assert np.array_equal(out, oracle.ntt(a, inverse=bool(i % 2)))
            d.free_buffer(b)
        with pytest.raises(zntt.ZpError):
            d.free_buffer(bufs[0])  # driver.cpp:160-170 double free
        with pytest.raises(zntt.ZpError):
            d.simple_evaluate(np.zeros((1 << 13) + 1, dtype=np.uint64))  # driver.cpp:289-297


def test_buffer_reenqueue_round_robin():
    """bench_throughput.cpp:76-86: each buffer enqueued several times without a
    wait in between, then one wait per evaluation.  Every evaluation restarts
    from the buffer's input, so a buffer ends holding the transform of its
    input in the direction of its last evaluation."""
    _say("ntt buffer re-enqueue")
    with zntt.NttGpuDriver(14, num_buffers=4) as d:
        bufs = [d.request_buffer() for _ in range(4)]
        with pytest.raises(zntt.ZpError):
            d.wait_for_result(bufs[0])  # never enqueued since the request
        inputs = [_rand(1 << 14, 90 + i) for i in range(4)]
        for b, a in zip(bufs, inputs):
            d.transfer_input_vector_to_user_buffer(b, a)
        last = {}
        for e in range(23):
            t, dr = e % 4, (e // 4) % 2
            d.enqueue_evaluation_async(bufs[t], dr)
            last[t] = dr
        for e in range(23):
            d.wait_for_result(bufs[e % 4])
        for t, b in enumerate(bufs):
            out = d.transfer_user_buffer_to_output_vector(b)
            assert np.array_equal(out, oracle.ntt(inputs[t], inverse=bool(last[t]))), t
            d.free_buffer(b)
        # a re-requested buffer starts over: waiting before an enqueue is an error again
        b = d.request_buffer()
        with pytest.raises(zntt.ZpError):
            d.wait_for_result(b)
        d.transfer_input_vector_to_user_buffer(b, inputs[0])
        d.enqueue_evaluation_async(b, zntt.FORWARD)
        d.wait_for_result(b)
        assert np.array_equal(d.transfer_user_buffer_to_output_vector(b), oracle.ntt(inputs[0]))
        d.free_buffer(b)


def test_batch_mixed_sizes():
    _say("ntt batch")
    rng = np.random.default_rng(4)
    log_ns = [int(x) for x in rng.integers(10, 17, size=12)]
    dirs = [int(x) for x in rng.integers(0, 2, size=12)]
    inputs = [_rand(1 << k, 70 + i) for i, k in enumerate(log_ns)]
    ptrs = []
    for a in inputs:
        p = ctypes.c_void_p()
        call("zp_device_malloc", ctypes.byref(p), a.nbytes)
        call("zp_memcpy_h2d", p, a.ctypes.data, a.nbytes)
        ptrs.append(p.value)
    try:
        zntt.ntt_batch_device(ptrs, log_ns, dirs)
        call("zp_device_synchronize")
        for p, a, k, dr in zip(ptrs, inputs, log_ns, dirs):
            out = np.empty_like(a)
            call("zp_memcpy_d2h", out.ctypes.data, ctypes.c_void_p(p), a.nbytes)
            assert np.array_equal(out, oracle.ntt(a, inverse=bool(dr))), (k, dr)
    finally:
        for p in ptrs:
            call("zp_device_free", ctypes.c_void_p(p))


def test_batch_graph_replays():
    """Graph-captured batch: every replay checked against the oracle, with
    eager launches and host synchronisations in between."""
    _say("ntt batch graph")
    rng = np.random.default_rng(8)
    log_ns = [10, 12, 13, 16, 16, 20, 24, 11, 18]
    dirs = [int(x) for x in rng.integers(0, 2, size=len(log_ns))]
    inputs = [_rand(1 << k, 300 + i) for i, k in enumerate(log_ns)]
    expect = [oracle.ntt(a, inverse=bool(d)) for a, d in zip(inputs, dirs)]
    ptrs = []
    for a in inputs:
        p = ctypes.c_void_p()
        call("zp_device_malloc", ctypes.byref(p), a.nbytes)
        ptrs.append(p.value)

    def upload():
        for p, a in zip(ptrs, inputs):
            call("zp_memcpy_h2d", ctypes.c_void_p(p), a.ctypes.data, a.nbytes)

    def check(tag):
        for p, a, e in zip(ptrs, inputs, expect):
            out = np.empty_like(a)
            call("zp_memcpy_d2h", out.ctypes.data, ctypes.c_void_p(p), a.nbytes)
            assert np.array_equal(out, e), tag

    b = zntt.NttBatch(ptrs, log_ns, dirs)
    try:
        for rep in range(3):
            upload()
            b.launch(use_graph=True)
            call("zp_device_synchronize")
            check(f"graph replay {rep}")
            upload()
            b.launch(use_graph=False)
            call("zp_device_synchronize")
            check(f"eager {rep}")
    finally:
        b.close()
        for p in ptrs:
            call("zp_device_free", ctypes.c_void_p(p))


@pytest.mark.parametrize("log_r", list(range(1, 12)))
def test_reverse_core_vs_oracle(log_r):
    """The REVERSE debug core (NttFpgaDriverArg::create_reverse) against the
    oracle's step-by-step restatement of reference.cpp:60-84."""
    _say(f"reverse core log_row_size={log_r}")
    x = _rand(1 << (2 * log_r), 300 + log_r)
    with zntt.NttGpuDriver.reverse(log_r) as d:
        assert d.input_vector_size() == 1 << (2 * log_r)
        assert np.array_equal(d.simple_evaluate(x), oracle.ntt_reverse(x, log_r))


def test_reverse_core_2p24_device_and_errors():
    """log_row_size 12 (2^24, the reference's size) through the device entry point,
    spot-checked with the element formula out[i][j] = in[R-1-j][R-1-i] w^(j (R-1-i));
    the per-pass timing works on the REVERSE plan; an inverse is rejected."""
    from hardcaml_zprize_amd.msm import DeviceBuffer

    _say("reverse core 2^24")
    log_r, R = 12, 1 << 12
    x = _rand(R * R, 77)
    dev = DeviceBuffer(x.nbytes).upload(x)
    try:
        with zntt.NttGpuDriver.reverse(log_r) as d:
            d.inplace_device(dev.ptr)
            call("zp_device_synchronize")
            got = dev.download(np.empty_like(x))
            w = pow(7, (P - 1) >> 24, P)
            rng = np.random.default_rng(5)
            for i, j in rng.integers(0, R, size=(2000, 2)):
                i, j = int(i), int(j)
                want = int(x[(R - 1 - j) * R + (R - 1 - i)]) * pow(w, j * (R - 1 - i), P) % P
                assert int(got[i * R + j]) == want
            m1, m2 = d.time_passes(dev.ptr, reps=2)
            assert m1 > 0 and m2 > 0
            with pytest.raises(zntt.ZpError):
                d.inplace_device(dev.ptr, zntt.INVERSE)
    finally:
        dev.free()
