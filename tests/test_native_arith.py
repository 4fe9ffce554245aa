"""CPU check of device-side arithmetic: the Goldilocks helpers (lazy
butterflies, power-of-two twiddles, 128-bit reduction, radix-64 register DFT)
and the Fr conversion used for the harness's data files.

The inline functions of hardcaml_zprize_amd/csrc/gl64.hpp and ntt4096.hpp are
compiled for the host with hipcc and compared with 128-bit modular arithmetic
and the defining DFT sum (tests/native/gl_lazy_check.cpp).  The GPU parity
tests then check that the kernels built on them match the oracle.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _build_and_run(tmp_path, name):
    exe = tmp_path / name
    src = os.path.join(ROOT, "tests", "native", name + ".cpp")
    inc = os.path.join(ROOT, "hardcaml_zprize_amd", "csrc")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-w", "-I", inc, src,
                    "-o", str(exe)], check=True, timeout=600)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_gl_lazy_arithmetic_host(tmp_path):
    _build_and_run(tmp_path, "gl_lazy_check")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fr_to_montgomery_host(tmp_path):
    """fr253.hpp (the harness-file scalar conversion s -> s 2^256 mod r) against
    shift-and-subtract big-integer arithmetic."""
    _build_and_run(tmp_path, "fr_check")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fq_binary_inverse_host(tmp_path):
    """fq377.hpp f_inv_fast (the device finalisation's inversion) against
    Fermat inversion and a * a^-1 = 1."""
    _build_and_run(tmp_path, "fq_inv_check")
