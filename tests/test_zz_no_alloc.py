"""trace_lbvh_rebuild_noalloc: the LBVH rebuild whose scratch is the library's
fixed device array (no hipMalloc), checked against trace_lbvh_rebuild.

Bar (integer and index work): sorted keys, permutation, canonical topology,
node boxes and closest hits bit for bit equal to the allocating rebuild's, on
two calls in a row (the second with a smaller mesh), eagerly and as replays of
a captured HIP graph on a context whose own scratch arena never grew."""
import ctypes as C

import numpy as np
import pytest

from trace_amd import KEY_M30, KEY_REF63, _lib, scenes
from trace_amd.device import Bvh, Context, Graph, Mesh

pytestmark = pytest.mark.gpu


def _noalloc(ctx, mesh, bvh, key_mode=KEY_REF63, enclosing=None):
    enc = None if enclosing is None else np.ascontiguousarray(enclosing, np.float64)
    return ctx.lib.trace_lbvh_rebuild_noalloc(ctx.h, mesh.h, None if enc is None else enc.ctypes.data_as(C.c_void_p),
                                              key_mode, bvh.h)


def _state(bvh, rays):
    keys, perm = bvh.sorted()
    rec, box = bvh.canonical()
    hit, t = bvh.closest_hit(rays)
    return keys, perm, rec, box, hit, t


def _assert_equal(a, b):
    for name, x, y in zip(("keys", "perm", "topology", "boxes", "hits", "t"), a, b):
        assert x.shape == y.shape, name
        assert np.array_equal(x, y), name


@pytest.fixture(scope="module")
def fresh():
    # a context whose scratch arena never grows: an allocating rebuild inside
    # a capture on it must fail, the no-allocation one must not
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("key_mode", [KEY_REF63, KEY_M30])
def test_noalloc_rebuild_matches_rebuild(ctx, fresh, key_mode):
    lib = ctx.lib
    rays = scenes.incoherent_rays(100_000, seed=43)
    for n, seed in ((200_000, 41), (50_000, 45)):  # the second call is smaller
        print(f"noalloc rebuild n={n} mode={key_mode}", flush=True)
        rest = scenes.random_triangles(n, seed=seed)
        moved = scenes.random_triangles(n, seed=seed + 1)
        m_rest, m_moved = Mesh(ctx, rest), Mesh(ctx, moved)
        ref = Bvh.lbvh(ctx, m_rest, key_mode=key_mode)
        ref.rebuild(m_moved, key_mode=key_mode)
        ctx.synchronize()
        want = _state(ref, rays)
        na = Bvh.lbvh(ctx, m_rest, key_mode=key_mode)
        rc = _noalloc(fresh, m_moved, na, key_mode)
        assert rc == _lib.TRACE_OK, lib.trace_last_error(fresh.h)
        fresh.synchronize()
        _assert_equal(_state(na, rays), want)
        # a caller box, passed without a host-memory copy
        box = np.array([-0.5, -0.25, -1.0, 1.5, 1.25, 2.0])
        ref.rebuild(m_moved, enclosing=box, key_mode=key_mode)
        ctx.synchronize()
        want_box = _state(ref, rays)
        assert _noalloc(fresh, m_moved, na, key_mode, enclosing=box) == _lib.TRACE_OK
        fresh.synchronize()
        _assert_equal(_state(na, rays), want_box)
        for b in (ref, na):
            b.close()
        m_rest.close()
        m_moved.close()


def test_noalloc_rebuild_in_a_graph(ctx, fresh):
    """Captured on the fresh context: each replay rebuilds from the moved
    positions, after eager no-allocation rebuilds from the rest positions
    changed the tree in between."""
    lib = ctx.lib
    print("noalloc rebuild in a graph", flush=True)
    n = 120_000
    rays = scenes.incoherent_rays(50_000, seed=47)
    rest = scenes.random_triangles(n, seed=48)
    moved = scenes.random_triangles(n, seed=49)
    m_rest, m_moved = Mesh(ctx, rest), Mesh(ctx, moved)
    ref = Bvh.lbvh(ctx, m_moved)
    want = _state(ref, rays)
    want_rest = _state(Bvh.lbvh(ctx, m_rest), rays)
    na = Bvh.lbvh(ctx, m_rest)
    # in the same capture, an allocating rebuild would need the fresh
    # context's arena to grow: it is refused (nothing enqueued)
    rcs = []

    def record():
        rcs.append(lib.trace_lbvh_rebuild(fresh.h, m_moved.h, None, KEY_REF63, na.h))
        fresh.check(_noalloc(fresh, m_moved, na))

    graph = Graph.capture(fresh, record)
    assert rcs == [_lib.TRACE_E_INVALID], rcs
    try:
        for k in range(3):
            assert _noalloc(fresh, m_rest, na) == _lib.TRACE_OK
            fresh.synchronize()
            _assert_equal(_state(na, rays), want_rest)
            graph.launch()
            fresh.synchronize()
            _assert_equal(_state(na, rays), want)
    finally:
        graph.close()


def test_noalloc_rebuild_refusals(ctx, fresh):
    lib = ctx.lib
    print("noalloc rebuild refusals", flush=True)
    # 3M triangles need more scratch than the fixed array holds: refused
    # before anything is enqueued, so the tree is unchanged
    pos = scenes.random_triangles(3_000_000, seed=50)
    mesh = Mesh(ctx, pos)
    bvh = Bvh.lbvh(ctx, mesh)
    keys0, perm0 = bvh.sorted()
    moved = Mesh(ctx, scenes.random_triangles(3_000_000, seed=51))
    assert _noalloc(fresh, moved, bvh) == _lib.TRACE_E_UNSUPPORTED
    assert b"scratch" in lib.trace_last_error(fresh.h)
    fresh.synchronize()
    keys1, perm1 = bvh.sorted()
    assert np.array_equal(keys0, keys1) and np.array_equal(perm0, perm1)
    # a mesh of another size than the tree's
    small = Mesh(ctx, scenes.random_triangles(1000, seed=52))
    assert _noalloc(fresh, small, bvh) == _lib.TRACE_E_INVALID
    assert _noalloc(fresh, small, bvh, key_mode=7) == _lib.TRACE_E_INVALID
    assert lib.trace_lbvh_rebuild_noalloc(None, small.h, None, 0, bvh.h) == _lib.TRACE_E_INVALID
