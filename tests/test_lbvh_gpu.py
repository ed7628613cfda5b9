"""GPU parity of the LBVH build against the FP64 oracle: Morton keys, sort
permutation, canonical topology and node boxes bit-exact; Cost() equal."""
import numpy as np
import pytest

from trace_amd import KEY_M30, KEY_REF63, scenes
from trace_amd.device import Bvh, Mesh

pytestmark = pytest.mark.gpu


def _build(ctx, pos, key_mode, enclosing=None):
    mesh = Mesh(ctx, pos)
    return mesh, Bvh.lbvh(ctx, mesh, enclosing=enclosing, key_mode=key_mode)


def _assert_same_tree(bvh, tree, pos):
    keys, perm = bvh.sorted()
    okeys, operm = tree.sorted()
    assert np.array_equal(keys, okeys), "Morton keys differ"
    # stable sort on both sides: the permutation is identical
    assert np.array_equal(perm, operm), "sort permutation differs"
    rec, box = bvh.canonical()
    orec, obox = tree.canonical()
    assert rec.shape == orec.shape, (rec.shape, orec.shape)
    assert np.array_equal(rec[:, 0], orec[:, 0])
    assert np.array_equal(rec[:, 1:3], orec[:, 1:3])
    # inputs are FP32-exact, so the FP32 boxes equal the FP64 boxes
    assert np.array_equal(box.astype(np.float64), obox)


@pytest.mark.parametrize("key_mode", [KEY_REF63, KEY_M30])
@pytest.mark.parametrize("n,seed", [(1, 1), (2, 1), (3, 2), (100, 3), (4095, 4), (4096, 5), (4097, 6), (50_000, 7)])
def test_lbvh_matches_oracle(ctx, oracle, key_mode, n, seed):
    print(f"lbvh parity n={n} seed={seed} mode={key_mode}", flush=True)
    pos = scenes.random_triangles(n, seed=seed)
    _, bvh = _build(ctx, pos, key_mode)
    tree = oracle.Tree(pos, key_mode=key_mode)
    _assert_same_tree(bvh, tree, pos)
    assert bvh.cost() == pytest.approx(tree.cost(), rel=1e-12)
    assert bvh.size() == tree.node_count()


@pytest.mark.parametrize("key_mode", [KEY_REF63, KEY_M30])
def test_lbvh_duplicates_collapse_into_leaves(ctx, oracle, key_mode):
    print(f"lbvh duplicates mode={key_mode}", flush=True)
    base = scenes.random_triangles(3000, seed=8)
    pos = np.concatenate([base, base[:1000], base[500:700], base[:10]])
    _, bvh = _build(ctx, pos, key_mode)
    _assert_same_tree(bvh, oracle.Tree(pos, key_mode=key_mode), pos)


def test_lbvh_caller_enclosing_box(ctx, oracle):
    print("lbvh caller enclosing", flush=True)
    pos = scenes.random_triangles(20_000, seed=9)
    enc = np.array([-0.5, -0.25, -1.0, 2.0, 1.5, 1.25])
    _, bvh = _build(ctx, pos, KEY_REF63, enclosing=enc)
    _assert_same_tree(bvh, oracle.Tree(pos, enclosing_box=enc), pos)


def test_lbvh_flat_scene_zero_extent_axis(ctx, oracle):
    print("lbvh flat scene", flush=True)
    pos = scenes.random_triangles(5000, seed=10)
    pos.reshape(-1, 3, 3)[:, :, 2] = 0.25  # every vertex in the plane z = 0.25
    _, bvh = _build(ctx, pos, KEY_REF63)
    _assert_same_tree(bvh, oracle.Tree(pos), pos)


def test_lbvh_cube_fixture(ctx, oracle):
    import os

    print("lbvh cube.obj", flush=True)
    pos, _ = scenes.parse_obj(os.path.join(os.path.dirname(__file__), "golden", "cube.obj"))
    _, bvh = _build(ctx, pos, KEY_REF63)
    _assert_same_tree(bvh, oracle.Tree(pos), pos)


@pytest.mark.parametrize("key_mode", [KEY_REF63, KEY_M30])
def test_lbvh_config2_full_size(ctx, oracle, key_mode):
    """1M triangles (seed 1): keys, permutation and topology bit-exact."""
    print(f"lbvh 1M mode={key_mode}", flush=True)
    pos = scenes.random_triangles(1_000_000, seed=1)
    _, bvh = _build(ctx, pos, key_mode)
    tree = oracle.Tree(pos, key_mode=key_mode)
    _assert_same_tree(bvh, tree, pos)
    info = bvh.info()
    assert info.depth <= 63


def test_rebuild_reuses_handle(ctx, oracle):
    print("lbvh rebuild", flush=True)
    a = scenes.random_triangles(30_000, seed=12)
    b = scenes.random_triangles(30_000, seed=13)
    mesh = Mesh(ctx, a)
    bvh = Bvh.lbvh(ctx, mesh)
    mesh.update(b)
    bvh.rebuild(mesh)
    ctx.synchronize()
    _assert_same_tree(bvh, oracle.Tree(b), b)


def test_refit_boxes_stable_across_rebuilds(ctx, oracle):
    """Regression test for the refit's cross-workgroup publication: the second
    child to arrive at a node must see its sibling's box.  Node boxes of eight
    rebuilds of the 1M-triangle LBVH all equal the oracle's."""
    print("lbvh refit stability 1M x 8", flush=True)
    pos = scenes.random_triangles(1_000_000, seed=1)
    mesh = Mesh(ctx, pos)
    bvh = Bvh.lbvh(ctx, mesh)
    _, obox = oracle.Tree(pos).canonical()
    for k in range(8):
        bvh.rebuild(mesh)
        _, box = bvh.canonical()
        assert np.array_equal(box.astype(np.float64), obox), f"rebuild {k}: node boxes differ"


def test_empty_and_oversized_meshes_are_rejected(ctx):
    """LBVH over zero primitives has no root (the reference's enclosing() of an
    empty slice is undefined): the boundary rejects it before any launch, and
    a malformed position array is refused rather than read out of bounds."""
    from trace_amd import _lib

    with pytest.raises(_lib.TraceError):
        Mesh(ctx, np.zeros((0, 9), np.float64))
    with pytest.raises((_lib.TraceError, ValueError)):
        Mesh(ctx, np.zeros((5, 8), np.float64))


@pytest.mark.parametrize("bits,kpt", [(8, 16), (8, 8), (8, 4), (10, 8), (10, 4), (11, 4)])
def test_every_sort_configuration_builds_the_same_tree(ctx, oracle, bits, kpt):
    """The key sort's digit width and tile size change the passes, not the
    result: keys, permutation and topology stay bit-exact for 63- and 30-bit
    keys, across tile edges, with duplicates, and in a forest (whose segment
    digits add 8-bit passes, here two of them for 300 segments)."""
    from trace_amd.device import set_tune

    print(f"sort configuration {bits} bits x {kpt} keys per thread", flush=True)
    set_tune(ctx, sort_bits=bits, sort_keys_per_thread=kpt)
    try:
        for key_mode in (KEY_REF63, KEY_M30):
            for n, seed in ((1, 1), (1023, 2), (1024, 3), (1025, 4), (70_000, 5)):
                pos = scenes.random_triangles(n, seed=seed)
                _, bvh = _build(ctx, pos, key_mode)
                _assert_same_tree(bvh, oracle.Tree(pos, key_mode=key_mode), pos)
            base = scenes.random_triangles(3000, seed=8)
            pos = np.concatenate([base, base[:1000], base[500:700]])
            _, bvh = _build(ctx, pos, key_mode)
            _assert_same_tree(bvh, oracle.Tree(pos, key_mode=key_mode), pos)
        rng = np.random.default_rng(5)
        parts = [scenes.random_triangles(int(k), seed=2000 + i) for i, k in enumerate(rng.integers(1, 40, 300))]
        forest = Bvh.lbvh(ctx, Mesh.segmented(ctx, parts))
        for i in (0, 255, 256, 299):
            v = forest.segment(i)
            _assert_same_tree(v, oracle.Tree(parts[i]), parts[i])
            v.close()
    finally:
        set_tune(ctx)


def test_autotune_build_keeps_the_fastest_sort(ctx, oracle):
    from trace_amd import _lib
    from trace_amd.device import SORT_CONFIGS, autotune_build, get_tune, set_tune

    print("autotune build", flush=True)
    pos = scenes.random_triangles(200_000, seed=12)
    mesh = Mesh(ctx, pos)
    try:
        ms = autotune_build(ctx, mesh, KEY_REF63)
        assert ms.shape == (len(SORT_CONFIGS),) and np.all(ms > 0)
        t = get_tune(ctx)
        assert (t.sort_bits, t.sort_keys_per_thread) == SORT_CONFIGS[int(np.argmin(ms))]
        _assert_same_tree(Bvh.lbvh(ctx, mesh), oracle.Tree(pos), pos)
        with pytest.raises(_lib.TraceError):
            set_tune(ctx, sort_bits=11, sort_keys_per_thread=16)  # 2048 counters per wave: tiles of 1024 only
        with pytest.raises(_lib.TraceError):
            set_tune(ctx, sort_bits=9)
    finally:
        set_tune(ctx)
