"""CPU check of the device-side Goldilocks arithmetic (lazy butterflies,
power-of-two twiddles, 128-bit reduction, radix-64 register DFT).

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


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_gl_lazy_arithmetic_host(tmp_path):
    exe = tmp_path / "gl_lazy_check"
    src = os.path.join(ROOT, "tests", "native", "gl_lazy_check.cpp")
    inc = os.path.join(ROOT, "hardcaml_zprize_amd", "csrc")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-w", "-I", inc, src,
                    "-o", str(exe)], check=True, timeout=600)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")
