"""The C++ NttGpuDriver mirror (include/zpgpu_ntt_driver.hpp) through the
host tool built beside the library."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "hardcaml_zprize_amd", "zp_ntt_host")


@pytest.mark.parametrize("args", [["test", "--log-n", "12", "--num-test-cases", "1"],
                                  ["test", "--log-n", "16", "--num-test-cases", "3"],
                                  ["bench", "--log-n", "14", "--num-evaluations", "10"]])
def test_host_tool(args):
    print("host tool", args, flush=True)
    r = subprocess.run([TOOL] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    if args[0] == "test":
        assert "FAILED" not in r.stdout and "OK" in r.stdout


def test_host_tool_rejects_unknown_flag():
    r = subprocess.run([TOOL, "test", "--bogus", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
