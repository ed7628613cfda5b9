"""Builds the gfx950 shared library ``trace_amd/libtrace_hip.so`` with hipcc,
and the C++ host API over it, ``trace_amd/libtrace_pt.so``, with g++.

Every ``.hip`` file under ``trace_amd/csrc`` is compiled to an object in
parallel and linked into one shared library exporting the C-ABI of
``include/trace_hip.h``.  ``trace_amd/pt/pt.cpp`` (the ``pkg/pt`` API of
``include/trace_pt.hpp``) is host code only; it links against
``libtrace_hip.so`` and finds it beside itself.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "trace_amd")
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD_DIR = os.path.join(PKG, "_build")
LIB_PATH = os.path.join(PKG, "libtrace_hip.so")
PT_SRC = os.path.join(PKG, "pt", "pt.cpp")
PT_LIB_PATH = os.path.join(PKG, "libtrace_pt.so")
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off (hipcc's host pass warns that it does
    # not know the feature; that warning is harmless)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compile all HIP sources and link ``libtrace_hip.so``; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")), 16)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH} from {len(srcs)} sources", flush=True)
    build_pt(verbose)
    return LIB_PATH


def build_pt(verbose: bool = True) -> str:
    """Compile the C++ host API into ``libtrace_pt.so`` (needs ``libtrace_hip.so``).
    FP64 host arithmetic is kept uncontracted, as the reference's Go code is."""
    tmp = PT_LIB_PATH + ".tmp"
    cmd = [CXX, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-ffp-contract=off", f"-I{INCLUDE}",
           PT_SRC, "-o", tmp, f"-L{PKG}", "-ltrace_hip", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {PT_SRC}:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, PT_LIB_PATH)
    if verbose:
        print(f"built {PT_LIB_PATH}", flush=True)
    return PT_LIB_PATH


if __name__ == "__main__":
    build()
    sys.exit(0)
