"""The C++ NttGpuDriver mirror (include/zpgpu_ntt_driver.hpp) through the
host tool built beside the library."""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "hardcaml_zprize_amd", "zp_ntt_host")


@pytest.mark.parametrize("args", [["test", "--log-n", "12", "--num-test-cases", "1"],
                                  ["test", "--log-n", "16", "--num-test-cases", "3"],
                                  ["bench", "--log-n", "14", "--num-evaluations", "10"],
                                  ["bench", "--log-n", "18", "--num-evaluations", "5", "--inverse"],
                                  ["throughput", "--log-n", "16", "--num-evaluations", "40"],
                                  ["throughput", "--log-n", "20", "--num-evaluations", "20", "--inverse"]])
def test_host_tool(args):
    print("host tool", args, flush=True)
    r = subprocess.run([TOOL] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    if args[0] == "test":  # each case also against the tool's CPU transform (test_random.cpp:56,73-90)
        assert "FAILED" not in r.stdout and "OK" in r.stdout and "mismatch" not in r.stdout
    if args[0] == "bench":  # every run compared with the CPU transform (bench_latency.cpp:95-98)
        assert "Incorrect result" not in r.stdout and "latency" in r.stdout
    if args[0] == "throughput":  # bench_throughput.cpp:94-103 output, every buffer checked
        assert "Incorrect result" not in r.stdout and "NTTs/second" in r.stdout


def test_host_tool_rejects_unknown_flag():
    r = subprocess.run([TOOL, "test", "--bogus", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


def _write_hex(path, vals, style):
    """'padded': one 0x%016x per line (the tool's output form); 'mixed': bare hex
    and 0x-prefixed tokens, several per line (what operator>> std::hex accepts)."""
    if style == "padded":
        text = "".join(f"0x{int(v):016x}\n" for v in vals)
    else:
        text = "".join((f"{int(v):x} " if i % 3 else f"0x{int(v):x}\n") for i, v in enumerate(vals))
    with open(path, "w") as f:
        f.write(text)


def _read_hex(path):
    return np.array([int(t, 16) for t in open(path).read().split()], dtype=np.uint64)


@pytest.mark.parametrize("core,log_n,inverse,layout", [("NTT", 12, False, "NORMAL_LAYOUT"),
                                                       ("NTT", 16, True, "OPTIMIZED_LAYOUT"),
                                                       ("NTT", 24, False, "OPTIMIZED_LAYOUT"),
                                                       ("REVERSE", 10, False, "NORMAL_LAYOUT")])
def test_evaluate_given_flow(tmp_path, core, log_n, inverse, layout):
    """evaluate_given.cpp: hex text in (with or without 0x), "0x%016x" lines out,
    checked against the oracle; --expected-filename reports PASSED / FAILED.  The
    FPGA's OPTIMIZED_LAYOUT (an internal device-buffer layout) gives the same
    natural-order vectors."""
    print("evaluate", core, log_n, inverse, flush=True)
    n = 1 << log_n
    rng = np.random.default_rng(log_n)
    x = rng.integers(0, oracle.P, size=n, dtype=np.uint64)
    inp, out, exp = tmp_path / "in.txt", tmp_path / "out.txt", tmp_path / "exp.txt"
    _write_hex(inp, x, "mixed")
    if core == "REVERSE":
        want = oracle.ntt_reverse(x, log_n // 2)
        size = ["--core-type", "REVERSE", "--log-row-size", str(log_n // 2)]
    else:
        want = oracle.ntt(x, inverse=inverse)
        size = ["--log-n", str(log_n)]
    _write_hex(exp, want, "padded")
    args = [TOOL, "evaluate"] + size + ["--input-filename", str(inp), "--output-filename", str(out),
                                       "--expected-filename", str(exp), "--xclbin", "unused",
                                       "--memory-layout", layout, "--log-blocks", "3"]
    if inverse:
        args.append("--inverse")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "EVALUATE PASSED" in r.stdout
    lines = open(out).read().splitlines()
    assert len(lines) == n and all(len(l) == 18 and l.startswith("0x") for l in lines)
    assert (_read_hex(out) == want).all()
    bad = want.copy()
    bad[n // 3] ^= np.uint64(1)
    _write_hex(exp, bad, "padded")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "EVALUATE FAILED" in r.stdout and f"i = {n // 3}" in r.stdout


def test_evaluate_input_count_errors(tmp_path):
    """evaluate_given.cpp:36-55: exactly n values, else an error naming the count."""
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    for count, msg in [(15, "but only got 15?"), (17, "but got more?")]:
        _write_hex(inp, np.arange(count, dtype=np.uint64), "padded")
        r = subprocess.run([TOOL, "evaluate", "--log-n", "4", "--input-filename", str(inp), "--output-filename",
                            str(out)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and msg in r.stderr
    r = subprocess.run([TOOL, "evaluate", "--log-n", "4", "--input-filename", str(tmp_path / "missing"),
                        "--output-filename", str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Failed to open input_file" in r.stderr


@pytest.mark.parametrize("args", [["test", "--core-type", "REVERSE", "--log-row-size", "6", "--num-test-cases", "2"],
                                  ["test", "--core-type", "REVERSE", "--log-row-size", "12", "--num-test-cases", "1"]])
def test_reverse_core_test_mode(args):
    print("host tool", args, flush=True)
    r = subprocess.run([TOOL] + args, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "NTT TEST PASSED" in r.stdout


def test_usage_errors():
    for args, msg in [(["test", "--core-type", "REVERSE"], "--log-row-size must be specified"),
                      (["test", "--memory-layout", "TILED"], "Unknown memory layout"),
                      (["test", "--core-type", "FFT"], "Unknown core type")]:
        r = subprocess.run([TOOL] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and msg in r.stderr
