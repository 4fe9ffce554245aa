"""The C-ABI library loads on a CPU-only host and exports every function
include/zpgpu.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from hardcaml_zprize_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "zpgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
        name = m.group(1)
        if name not in ("if", "sizeof"):
            names.add(name)
    return sorted(names)


def test_header_parsed():
    names = declared_functions()
    assert "msm_init" in names and "msm_mult" in names and "zp_ntt_create" in names
    assert len(names) > 40


def test_every_declared_symbol_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_loader_signatures_cover_header():
    assert sorted(_lib.SIGNATURES) == declared_functions()
    assert _lib.missing_symbols() == []


def test_status_and_version_without_gpu():
    lib = _lib.lib()
    assert lib.zp_version().decode().startswith("zpgpu")
    assert lib.zp_last_error() is not None


def test_reference_symbol_names_present():
    """msm_init / msm_mult are the exact symbols test_fpga_harness/src/lib.rs:19-33 binds."""
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in ("msm_init", "msm_mult"):
        assert hasattr(lib, n)
