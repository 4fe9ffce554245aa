"""The host-side code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only): the library's option parsing, window layouts and cost model
(msm_plan.hpp), its field arithmetic and inversion as used by the MSM
finalisation (fq377.hpp, msm_host.hpp), the Goldilocks helpers of the NTT
kernels (gl64.hpp, ntt4096.hpp), the harness-file scalar conversion
(fr253.hpp), and the oracle's C code.  Each program checks results against
an independent property and must exit 0 with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
CSRC = os.path.join(ROOT, "hardcaml_zprize_amd", "csrc")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def _have(tool):
    return shutil.which(tool) is not None and os.path.exists("/opt/rocm/include/hip/hip_runtime.h")


def _run(exe, *args, timeout=600):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout, env=ENV)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, \
        r.stdout[-2000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("ok")


def _build_cpp(tmp_path, name):
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", CSRC, *SAN,
                    os.path.join(NATIVE, name + ".cpp"), "-o", str(exe)], check=True, timeout=600)
    return exe


@pytest.mark.skipif(not _have("g++"), reason="g++ or HIP headers missing")
@pytest.mark.parametrize("name,args", [("host_plan_check", [20000]), ("fq_inv_check", [1000]), ("fr_check", [])])
def test_library_host_code_sanitized(tmp_path, name, args):
    _run(_build_cpp(tmp_path, name), *args)


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc missing")
def test_ntt_host_arithmetic_sanitized(tmp_path):
    """gl_lazy_check includes the NTT kernel headers, so it is built by hipcc;
    the sanitizers apply to its host code only."""
    exe = tmp_path / "gl_lazy_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O1", "-g",
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-I", CSRC,
                    os.path.join(NATIVE, "gl_lazy_check.cpp"), "-o", str(exe)], check=True, timeout=600)
    env = dict(ENV, ASAN_OPTIONS="detect_leaks=0")  # the HIP runtime keeps allocations until exit
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "runtime error" not in r.stderr, r.stdout[-2000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("ok")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
def test_oracle_sanitized(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    subprocess.run(["gcc", "-std=c11", "-pthread", *SAN, "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(NATIVE, "oracle_sanitize.c"), os.path.join(ROOT, "oracle", "goldilocks.c"),
                    os.path.join(ROOT, "oracle", "bls12_377.c"), "-o", str(exe)], check=True, timeout=600)
    _run(exe)
