"""Runs the native C++ API test program (tests/cpp/test_api_gpu.cpp), built
by darcy-cpp_amd/build.py against libdarcy_hip.so, on the GPU."""

from __future__ import annotations

import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BINARY = os.path.join(ROOT, "tests", "cpp", "test_api_gpu")


def test_cpp_api_program():
    assert os.path.exists(BINARY), "build first: python darcy-cpp_amd/build.py"
    print("running tests/cpp/test_api_gpu", flush=True)
    proc = subprocess.run([BINARY], capture_output=True, text=True, timeout=120)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert " 0 failures" in proc.stdout
