"""The NTT tile table (zp_ntt_set_tiles / zp_ntt_get_tiles) on the host: the
validation of a choice, the built-in policy, and the autotuner's JSON table.
These calls only read and write the library's process-wide table; nothing
runs on a GPU."""
import ctypes
import json

import pytest

from hardcaml_zprize_amd import autotune
from hardcaml_zprize_amd import ntt as zntt
from hardcaml_zprize_amd._lib import ZpError, call


@pytest.fixture(autouse=True)
def _builtin_policy():
    yield
    for lg in range(zntt.TILE_MIN_LOG, zntt.TILE_MAX_LOG + 1):
        for b in (False, True):
            zntt.set_tiles(lg, b)


def builtin(log_n):
    """The library's built-in choice restated: balanced split (4096-point
    columns from 2^21), sequences per workgroup up to 2^13 LDS elements and few
    enough to keep 2^8 workgroups."""
    n1 = 12 if log_n > 20 else (log_n + 1) // 2
    n2 = log_n - n1

    def t(log_l, log_seq):
        lt = max(0, 13 - log_l)
        lt = min(lt, max(0, log_seq - 8))
        return min(lt, log_seq)

    return n1, t(n1, n2), t(n2, n1)


@pytest.mark.parametrize("log_n", range(10, 24))
def test_builtin_policy(log_n):
    for b in (False, True):
        assert zntt.get_tiles(log_n, b) == builtin(log_n)


def test_candidates_are_exactly_the_accepted_choices():
    for lg in (10, 15, 20, 22):
        cands = set(zntt.tile_candidates(lg))
        for n1 in range(0, 14):
            for t1 in range(0, 15):
                for t2 in range(0, 15):
                    try:
                        zntt.set_tiles(lg, False, n1, t1, t2)
                        ok = True
                    except ZpError:
                        ok = False
                    assert ok == ((n1, t1, t2) in cands), (lg, n1, t1, t2)
        assert builtin(lg) in cands


def test_set_get_and_modes_are_separate():
    zntt.set_tiles(16, True, 6, 3, 1)
    assert zntt.get_tiles(16, True) == (6, 3, 1)
    assert zntt.get_tiles(16, False) == builtin(16)
    zntt.set_tiles(16, False, 9, -1, -1)  # split only: tiles from the built-in rule for that split
    n1, t1, t2 = zntt.get_tiles(16, False)
    # 2^7 columns of 2^9: one per workgroup (fewer than 2^8 sequences);
    # 2^9 rows of 2^7: two per workgroup keep 2^8 workgroups
    assert (n1, t1, t2) == (9, 0, 1)
    zntt.set_tiles(16, False)
    assert zntt.get_tiles(16, False) == builtin(16)


@pytest.mark.parametrize("args", [(9, 0, -1, -1, -1), (16, 2, -1, -1, -1),
                                  (16, 0, -2, -1, -1), (16, 0, 13, -1, -1), (16, 0, 3, -1, -1),
                                  (16, 0, 8, 6, -1), (16, 0, 8, -1, 6), (20, 1, 8, -1, 6),
                                  (21, 0, 8, -1, -1), (22, 1, 12, 2, -1), (23, 0, -1, -1, 3), (24, 0, -1, -1, -1)])
def test_invalid_choices_rejected(args):
    with pytest.raises(ZpError):
        call("zp_ntt_set_tiles", *args)
    with pytest.raises(ZpError):
        call("zp_ntt_get_tiles", 9, 0, ctypes.byref(ctypes.c_int()), ctypes.byref(ctypes.c_int()),
             ctypes.byref(ctypes.c_int()))


def test_autotune_table_round_trip(tmp_path):
    path = tmp_path / "t.json"
    doc = {"msm_window": {"12": 8}, "ntt_tiles": {"single": {"12": [6, 1, 1]}, "batched": {"14": [7, 2, 3]}}}
    path.write_text(json.dumps(doc))
    assert autotune.load(str(path)) == {12: 8}
    tiles = autotune.load_ntt_tiles(str(path))
    assert tiles == {"single": {12: (6, 1, 1)}, "batched": {14: (7, 2, 3)}}
    assert autotune.apply_ntt_tiles(path=str(path)) == 2
    assert zntt.get_tiles(12, False) == (6, 1, 1)
    assert zntt.get_tiles(14, True) == (7, 2, 3)
    assert zntt.get_tiles(12, True) == builtin(12)
    assert autotune.load_ntt_tiles(str(tmp_path / "absent.json")) == {}


def test_committed_table_is_valid():
    """Every committed tile choice is one the library accepts."""
    for mode, per in autotune.load_ntt_tiles().items():
        for lg, choice in per.items():
            assert choice in zntt.tile_candidates(lg), (mode, lg, choice)
