"""Host build of the vertical walks (pre-smoothing from zero and residual +
restriction, darcy-cpp_amd/csrc/device.hpp) against plain loops, bit for
bit, under AddressSanitizer (tests/walk_check.hip): whole levels as the grid
path runs them, the persistent kernel's row-pipelined fresh down visit step
by step, and the FP32 cycle's single-colour walks (red from final black
cells; black from final red cells, with and without the prolongation)."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "darcy-cpp_amd", "csrc")


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_walks_match_plain_loops(tmp_path):
    exe = str(tmp_path / "walk_check")
    cmd = ["hipcc", "-std=c++17", "-O1", "-ffp-contract=off", "--offload-arch=gfx950",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
           "-I", CSRC, os.path.join(HERE, "walk_check.hip"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    lines = out.stdout.strip().splitlines()
    assert lines[-1] == "ALL OK"
    assert sum(1 for l in lines if l.endswith(": ok")) == 1880
