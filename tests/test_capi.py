"""CPU checks of the C-ABI library: it loads, exports every symbol the headers
declare, and its error paths need no device."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(trace_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ["trace_ctx_create", "trace_lbvh_build", "trace_closest_hit", "trace_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from trace_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers every declaration
    assert set(declared_symbols()) <= set(_lib.SIGNATURES), set(declared_symbols()) - set(_lib.SIGNATURES)


def test_error_paths_without_device():
    from trace_amd import _lib

    lib = _lib.load()
    assert lib.trace_version() >= 1
    # null handles are rejected without touching a device
    assert lib.trace_ctx_destroy(None) == _lib.TRACE_E_INVALID
    assert lib.trace_mesh_destroy(None) == _lib.TRACE_E_INVALID
    assert lib.trace_bvh_destroy(None) == _lib.TRACE_E_INVALID
    assert lib.trace_ctx_create(0, None) == _lib.TRACE_E_INVALID
    assert lib.trace_last_error(None) == b"null context"


def test_no_device_here_is_reported():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from trace_amd import _lib

    lib = _lib.load()
    h = C.c_void_p()
    assert lib.trace_ctx_create(0, C.byref(h)) == _lib.TRACE_E_NODEVICE
