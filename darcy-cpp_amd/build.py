"""Build the darcy GPU library and its Python extension in-tree.

    python darcy-cpp_amd/build.py

Produces, next to the Python package (darcy-cpp_amd/darcy/):
  libdarcy_hip.so                    HIP kernels for gfx950 + the C ABI (include/darcy_hip.h)
  _core.cpython-*.so                 pybind11 module: C++17 API + bindings, linked to the library

Objects go to darcy-cpp_amd/_build/.  Every HIP translation unit is compiled
with -ffp-contract=off (the FP64 path reproduces the reference's rounding)
and with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "darcy")
OBJ = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("DARCY_OFFLOAD_ARCH", "gfx950")
VERSION = "1.0.0"

HIP_SOURCES = ["batch.hip", "grid.hip", "conv.hip", "fields.hip", "capi.cpp", "net.cpp"]
HOST_SOURCES = ["host/api.cpp", "bindings.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    noise = ("packed-fp32-ops",)
    lines = [ln for ln in proc.stdout.splitlines() if not any(t in ln for t in noise)]
    if proc.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + "\n".join(lines))
    if lines and os.environ.get("DARCY_BUILD_VERBOSE"):
        print("\n".join(lines))


def _hip_object(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run(["hipcc", *HIPCC_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def _host_object(src: str) -> str:
    import pybind11

    obj = os.path.join(OBJ, os.path.basename(src) + ".host.o")
    _run([
        "g++", "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", f"-DDARCY_VERSION=\"{VERSION}\"",
        f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
        "-c", os.path.join(CSRC, src), "-o", obj,
    ])
    return obj


def build(verbose: bool = False) -> dict[str, str]:
    """Compile everything; returns the paths of the two shared objects."""
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        hip_objs = list(pool.map(_hip_object, HIP_SOURCES))
        host_objs = list(pool.map(_host_object, HOST_SOURCES))
    lib = os.path.join(OUT, "libdarcy_hip.so")
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *hip_objs, "-o", lib])
    ext = os.path.join(OUT, "_core" + sysconfig.get_config_var("EXT_SUFFIX"))
    _run(["g++", "-shared", "-fPIC", "-fopenmp", *host_objs, "-o", ext, f"-L{OUT}", "-ldarcy_hip",
          "-Wl,-rpath,$ORIGIN"])
    # The native test of the C++ API: the API's object linked into a program
    # against the library (the way a C++ caller uses it).
    cpp_test = os.path.join(ROOT, "tests", "cpp", "test_api_gpu")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", os.path.join(ROOT, "tests", "cpp", "test_api_gpu.cpp"),
          host_objs[0], "-o", cpp_test, "-fopenmp", f"-L{OUT}", "-ldarcy_hip",
          "-Wl,-rpath,$ORIGIN/../../darcy-cpp_amd/darcy"])
    if verbose:
        print(f"built {lib}\nbuilt {ext}\nbuilt {cpp_test}")
    return {"lib": lib, "ext": ext, "cpp_test": cpp_test}


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
