"""Builds ``tests/cpp/_build/pt_test``: the C++ tests of ``include/trace_pt.hpp``,
linked against ``trace_amd/libtrace_pt.so``, ``trace_amd/libtrace_hip.so`` and
the oracle (``oracle/_build/liboracle.so``, the checker).  Host code only (g++);
the libraries are found relative to the binary, so the tree can move."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "pt_test.cpp")
OUT_DIR = os.path.join(HERE, "_build")
BINARY = os.path.join(OUT_DIR, "pt_test")


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    pkg = os.path.join(ROOT, "trace_amd")
    orc = os.path.join(ROOT, "oracle", "_build")
    tmp = BINARY + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}", SRC,
           "-o", tmp, f"-L{pkg}", "-ltrace_pt", "-ltrace_hip", f"-L{orc}", "-loracle",
           "-Wl,-rpath,$ORIGIN/../../../trace_amd:$ORIGIN/../../../oracle/_build"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {SRC}:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, BINARY)
    if verbose:
        print(f"built {BINARY}", flush=True)
    return BINARY


if __name__ == "__main__":
    build()
