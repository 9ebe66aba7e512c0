"""Builds the in-tree native libraries for gfx950.

  sors_amd/libsors_amd.so   HIP kernels + C ABI (include/sors_amd.h)
  sors_amd/py_sors*.so      pybind11 module mirroring the reference's py_sors
  tests/cpp/test_facade     C++ test of the host interfaces (run by a GPU test)

Run ``python -m sors_amd.build`` (or ``__graft_entry__.build()``).  The build
always recompiles; nothing is cached outside the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CPP = os.path.join(PKG, "cpp")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsors_amd.so")
ARCH = os.environ.get("SORS_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # hardware float atomics for the CSR scatter (coarse-grained hipMalloc memory)
    "-munsafe-fp-atomics",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-add fusion decided per source expression, so a device function
    # rounds the same way in every kernel that inlines it (a batched launch
    # and a single-rollout launch give identical results)
    "-ffp-contract=on",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    os.makedirs(BUILD, exist_ok=True)
    host_o = []
    for name in ("pattern", "coarse"):
        o = os.path.join(BUILD, name + ".o")
        _run([CXX, "-std=c++17", "-O2", "-fPIC", "-pthread", "-I", CSRC, "-c", os.path.join(CSRC, name + ".cpp"),
              "-o", o])
        host_o.append(o)
    hip_o = os.path.join(BUILD, "sors_amd.o")
    _run([HIPCC, *HIP_FLAGS, "-I", INC, "-I", CSRC, "-c", os.path.join(CSRC, "sors_amd.hip"), "-o", hip_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, hip_o, *host_o, "-pthread",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_pybind() -> str | None:
    src = os.path.join(CPP, "py_sors.cpp")
    if not os.path.exists(src):
        return None
    try:
        import pybind11
    except ImportError:  # pragma: no cover - pybind11 is part of the image
        print("pybind11 not importable; skipping py_sors", file=sys.stderr)
        return None
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "py_sors" + suffix)
    srcs = [os.path.join(CPP, f) for f in sorted(os.listdir(CPP)) if f.endswith(".cpp")]
    cmd = [CXX, "-std=c++17", "-O2", "-fPIC", "-shared", "-I", INC, "-I", CPP, "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"], *srcs, "-L", PKG, "-lsors_amd", "-Wl,-rpath,$ORIGIN",
           "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_cpp_tests() -> str | None:
    """The C++ test program of the host interfaces, linked against the library."""
    src = os.path.join(ROOT, "tests", "cpp", "test_facade.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "test_facade")
    srcs = [src] + [os.path.join(CPP, f) for f in ("sors_elements.cpp", "sors_facade.cpp", "sors_io.cpp")]
    _run([CXX, "-std=c++17", "-O2", "-I", INC, "-I", CPP, *srcs, "-L", PKG, "-lsors_amd",
          "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/../../sors_amd", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build() -> None:
    build_lib()
    build_pybind()
    build_cpp_tests()


if __name__ == "__main__":
    build()
