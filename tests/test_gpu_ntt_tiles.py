"""Every tile choice of the two-pass LDS transform (zp_ntt_set_tiles: the
column length and the sequences per workgroup of each pass) gives the
oracle's output, forward and inverse, for single transforms and for batch
groups (jobs on grid.y, replayed from a graph)."""
import sys

import numpy as np
import pytest

import oracle
from hardcaml_zprize_amd import msm as zmsm
from hardcaml_zprize_amd import ntt as zntt
from hardcaml_zprize_amd._lib import call

pytestmark = pytest.mark.gpu
P = oracle.P


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _builtin_policy():
    yield
    for lg in range(zntt.TILE_MIN_LOG, zntt.TILE_MAX_LOG + 1):
        for b in (False, True):
            zntt.set_tiles(lg, b)


def _rand(n, seed):
    return np.random.default_rng(seed).integers(0, P, size=n, dtype=np.uint64)


@pytest.mark.parametrize("log_n", [10, 13, 16, 20, 21, 22, 23])
def test_every_tile_single(log_n):
    a = _rand(1 << log_n, 500 + log_n)
    fwd, inv = oracle.ntt(a), oracle.ntt(a, inverse=True)
    cands = zntt.tile_candidates(log_n)
    _say(f"ntt tiles single log_n={log_n}: {len(cands)} choices")
    with zntt.NttGpuDriver(log_n) as d:
        for c in cands:
            zntt.set_tiles(log_n, False, *c)
            assert zntt.get_tiles(log_n, False) == c
            assert np.array_equal(d.simple_evaluate(a, zntt.FORWARD), fwd), c
            assert np.array_equal(d.simple_evaluate(a, zntt.INVERSE), inv), c


@pytest.mark.parametrize("log_n,stride", [(10, 1), (14, 3), (17, 2), (20, 1), (22, 1)])
def test_tiles_batched(log_n, stride):
    """Three jobs of one size (two forward, one inverse) per choice; the graph
    is captured after the choice is set and replayed twice."""
    xs = [_rand(1 << log_n, 700 + 3 * log_n + i) for i in range(3)]
    dirs = [0, 1, 0]
    want = [oracle.ntt(x, inverse=bool(d)) for x, d in zip(xs, dirs)]
    bufs = [zmsm.DeviceBuffer(8 << log_n) for _ in xs]
    cands = zntt.tile_candidates(log_n)[::stride]
    _say(f"ntt tiles batched log_n={log_n}: {len(cands)} choices")
    try:
        for c in cands:
            zntt.set_tiles(log_n, True, *c)
            b = zntt.NttBatch([x.ptr for x in bufs], [log_n] * 3, dirs)
            try:
                for _ in range(2):
                    for buf, x in zip(bufs, xs):
                        buf.upload(x)
                    b.launch(use_graph=True)
                    call("zp_device_synchronize")
                    for buf, w in zip(bufs, want):
                        assert np.array_equal(buf.download(np.empty(1 << log_n, dtype=np.uint64)), w), c
            finally:
                b.close()
    finally:
        for buf in bufs:
            buf.free()


def test_autotuner_ntt_leaves_valid_applied_choices():
    """The tuner's NTT search on two size classes: every chosen tile is a
    valid one, it is what the library then uses, and a transform under it
    still equals the oracle."""
    from hardcaml_zprize_amd import autotune

    _say("ntt autotuner search 2^11, 2^15")
    res = autotune.tune_ntt([11, 15], batch_jobs=4)
    for mode, per in res.items():
        for lg, (choice, times) in per.items():
            assert choice in zntt.tile_candidates(lg)
            assert zntt.get_tiles(lg, mode == "batched") == choice
            assert "%d,%d,%d" % choice in times and len(times) >= 3
    a = _rand(1 << 15, 99)
    with zntt.NttGpuDriver(15) as d:
        assert np.array_equal(d.simple_evaluate(a, zntt.FORWARD), oracle.ntt(a))


@pytest.mark.parametrize("batched", [False, True])
def test_committed_table_every_size(batched):
    """The committed autotuner table (what MixedBatch, the server and the bench
    apply): every size class 2^10 .. 2^23, forward and inverse, equals the
    oracle; batched runs two jobs of the size on grid.y."""
    from hardcaml_zprize_amd import autotune

    tiles = autotune.load_ntt_tiles().get("batched" if batched else "single", {})
    assert sorted(tiles) == list(range(zntt.TILE_MIN_LOG, zntt.TILE_MAX_LOG + 1))
    autotune.apply_ntt_tiles({"batched" if batched else "single": tiles})
    for lg in sorted(tiles):
        assert zntt.get_tiles(lg, batched) == tiles[lg]
        _say(f"ntt committed tiles log_n={lg} batched={batched}: {tiles[lg]}")
        xs = [_rand(1 << lg, 900 + 2 * lg + i) for i in range(2)]
        if not batched:
            with zntt.NttGpuDriver(lg) as d:
                assert np.array_equal(d.simple_evaluate(xs[0], zntt.FORWARD), oracle.ntt(xs[0]))
                assert np.array_equal(d.simple_evaluate(xs[1], zntt.INVERSE), oracle.ntt(xs[1], inverse=True))
            continue
        bufs = [zmsm.DeviceBuffer(8 << lg).upload(x) for x in xs]
        try:
            b = zntt.NttBatch([x.ptr for x in bufs], [lg, lg], [0, 1])
            try:
                b.launch(use_graph=True)
                call("zp_device_synchronize")
            finally:
                b.close()
            got = [buf.download(np.empty(1 << lg, dtype=np.uint64)) for buf in bufs]
            assert np.array_equal(got[0], oracle.ntt(xs[0])), lg
            assert np.array_equal(got[1], oracle.ntt(xs[1], inverse=True)), lg
        finally:
            for buf in bufs:
                buf.free()
