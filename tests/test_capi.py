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
    rc = lib.trace_ctx_create(0, C.byref(h))
    if rc == _lib.TRACE_OK:  # the library found a device that torch does not report
        lib.trace_ctx_destroy(h)
        pytest.skip("a device is present")
    assert rc == _lib.TRACE_E_NODEVICE


@pytest.mark.gpu
def test_oversized_calls_are_refused_before_launch(ctx):
    """Sizes the 32-bit ray counter and int32 queues cannot index are refused
    with TRACE_E_UNSUPPORTED before any launch: the pointers below are never
    dereferenced (a launch with them would fault)."""
    import numpy as np

    from trace_amd import _lib, scenes
    from trace_amd.device import Bvh, Mesh

    bvh = Bvh.lbvh(ctx, Mesh(ctx, scenes.random_triangles(100, seed=1)))
    dummy = C.c_void_p(0x1000)
    for n in (2**31, 2**32, 2**32 + 5):
        rc = ctx.lib.trace_closest_hit_device(ctx.h, bvh.h, dummy, n, dummy, dummy)
        assert rc == _lib.TRACE_E_UNSUPPORTED, (n, rc)
    parts = [scenes.random_triangles(50, seed=s) for s in (2, 3)]
    mesh = Mesh.segmented(ctx, parts)
    mesh.set_shading(np.zeros(mesh.ntri, np.int32), np.array([[0, 0.5, 0.5, 0.5]], np.float32))
    forest = Bvh.lbvh(ctx, mesh)
    from trace_amd.device import camera_struct, render_params

    cam = camera_struct(scenes.camera_cfg2())
    segs = np.zeros(8, np.int32)
    cams = (_lib.Camera * 8)(*([cam] * 8))
    # 8 scenes of 16384 x 8193 pixels: each image is under 2^28 pixels, one
    # sample of all eight is over 2^30 paths
    p = render_params(16384, 8193, 1, 1, 0, 1)
    rc = ctx.lib.trace_render_batched(ctx.h, mesh.h, forest.h, 8, segs.ctypes.data_as(C.c_void_p), cams,
                                      C.byref(p), dummy, dummy, None)
    assert rc == _lib.TRACE_E_UNSUPPORTED, rc
    p = render_params(32768, 8193, 1, 1, 0, 1)  # one image over 2^28 pixels
    rc = ctx.lib.trace_render_batched(ctx.h, mesh.h, forest.h, 1, segs.ctypes.data_as(C.c_void_p), cams,
                                      C.byref(p), dummy, dummy, None)
    assert rc == _lib.TRACE_E_UNSUPPORTED, rc
