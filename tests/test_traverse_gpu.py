"""GPU parity of closest-hit traversal against the FP64 oracle.

Bar: hit triangle IDs agree on >= 99.99 % of rays; where both hit the same
triangle, t agrees within 1e-5 relative.  Mismatches can only come from FP32
vs FP64 rounding at triangle edges or the |det| <= 1e-7 boundary."""
import math

import numpy as np
import pytest

from trace_amd import KEY_M30, KEY_REF63, scenes
from trace_amd.device import Bvh, Mesh, set_tune

pytestmark = pytest.mark.gpu

HIT_AGREE = 0.9999
T_REL = 1e-5


def _check(hit, t, ohit, ot, min_hits=0.1):
    agree = np.mean(hit == ohit)
    assert agree >= HIT_AGREE, f"hit agreement {agree:.6f}"
    both = (hit == ohit) & (hit >= 0)
    rel = np.abs(t[both].astype(np.float64) - ot[both]) / np.maximum(np.abs(ot[both]), 1e-30)
    assert rel.size == 0 or rel.max() <= T_REL, f"max rel t error {rel.max():.3g}"
    miss = hit < 0
    assert np.all(np.isinf(t[miss]) | (t[miss] > 0))
    assert np.mean(hit >= 0) >= min_hits


@pytest.mark.parametrize("key_mode", [KEY_REF63, KEY_M30])
def test_incoherent_small(ctx, oracle, key_mode):
    print(f"traverse incoherent 20k tris mode={key_mode}", flush=True)
    pos = scenes.random_triangles(20_000, seed=21, mean_edge=0.02)
    rays = scenes.incoherent_rays(50_000, seed=22)
    mesh = Mesh(ctx, pos)
    bvh = Bvh.lbvh(ctx, mesh, key_mode=key_mode)
    hit, t = bvh.closest_hit(rays)
    ohit, ot = oracle.Tree(pos, key_mode=key_mode).closest_hit(rays)
    _check(hit, t, ohit, ot)


def test_against_brute_force(ctx, oracle):
    print("traverse vs brute force", flush=True)
    pos = scenes.random_triangles(2000, seed=23, mean_edge=0.05)
    rays = scenes.incoherent_rays(4000, seed=24)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    hit, t = bvh.closest_hit(rays)
    ohit, ot = oracle.closest_hit_brute(pos, rays)
    _check(hit, t, ohit, ot)


def test_edge_cases(ctx):
    print("traverse edge cases", flush=True)
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, -1, 1, 0, -1, 0, 1, -1]], np.float64)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, tri))
    inf = math.inf
    rays = np.array([
        [0.25, 0.25, 2.0, 0.001, 0, 0, -1, inf],    # hits z=0 at t=2
        [0.25, 0.25, 2.0, 0.001, 0, 0, -4, inf],    # unnormalised: t=0.5
        [0.75, 0.75, 2.0, 0.001, 0, 0, -1, inf],    # outside both
        [0.25, 0.25, 2.0, 0.001, 0, 0, -1, 2.0],    # t == tmax accepted
        [0.25, 0.25, 2.0, 0.001, 0, 0, -1, 1.999],  # t > tmax rejected
        [0.25, 0.25, -2.0, 0.001, 0, 0, -1, inf],   # behind
        [0.25, 0.25, -2.0, 0.001, 0, 0, 1, inf],    # from below: z=-1 first, t=1
        [0.25, 0.25, 2.0, 0.001, 1, 0, 0, inf],     # parallel to the plane
    ], np.float32)
    hit, t = bvh.closest_hit(rays)
    assert hit.tolist() == [0, 0, -1, 0, -1, -1, 1, -1]
    assert t[0] == 2.0 and t[1] == 0.5 and t[6] == 1.0
    assert np.isinf(t[2]) and t[4] == np.float32(1.999)


def test_tie_goes_to_reference_winner(ctx, oracle):
    print("traverse tie rule", flush=True)
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    pos = np.array([tri, tri, tri], np.float64)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    ray = np.array([[0.25, 0.25, 1.0, 0.001, 0, 0, -1, math.inf]], np.float32)
    hit, _ = bvh.closest_hit(ray)
    ohit, _ = oracle.Tree(pos).closest_hit(ray)
    assert hit[0] == ohit[0] == 2  # last in sorted (stable) order


@pytest.mark.parametrize("threads", [64, 256])
def test_ray_counts_around_the_claim_size(ctx, oracle, threads):
    # The persistent kernel hands rays to its waves in chunks of 64 consecutive
    # indices.  The full batch (a ragged last chunk) meets the oracle bar, and
    # every prefix shorter than, equal to or just past one or two chunks gives
    # bit for bit the full batch's results for its rays.
    print("traverse ray counts around the chunk size", threads, flush=True)
    pos = scenes.random_triangles(3000, seed=31)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    rays = scenes.incoherent_rays(4097, seed=32)
    set_tune(ctx, trav_threads=threads, trav_persistent=True)
    try:
        full_hit, full_t = bvh.closest_hit(rays)
        ohit, ot = oracle.Tree(pos).closest_hit(rays)
        _check(full_hit, full_t, ohit, ot, min_hits=0.0)
        for n in (1, 63, 64, 65, 127, 128, 129):
            hit, t = bvh.closest_hit(rays[:n])
            assert np.array_equal(hit, full_hit[:n]), n
            assert np.array_equal(t, full_t[:n]), n
    finally:
        set_tune(ctx)


def test_origin_order_matches_input_order(ctx, oracle):
    # Batches of 2^18 rays or more are traced grouped by origin cell (64 cells
    # of the root box, copied 1024 rays per workgroup), unless one cell holds
    # more than half of them (a camera's rays: traced in input order).  Results
    # are each ray's own, bit for bit the input-order trace's: on the LBVH and
    # the PHR tree, for a ragged batch with origins far outside the root box and
    # on its faces, for prefixes just below, at and above the size where
    # grouping starts, for a camera batch and for a batch that is 60 % one cell.
    print("traverse origin order vs input order", flush=True)
    pos = scenes.random_triangles(100_000, seed=41, mean_edge=0.01)
    rays = scenes.incoherent_rays(300_001, seed=42)
    rays[:2000, 0:3] = rays[:2000, 0:3] * 20.0 - 10.0
    lo, hi = pos.reshape(-1, 3).min(0), pos.reshape(-1, 3).max(0)
    rays[2000:2100, 0:3] = hi.astype(np.float32)
    rays[2100:2200, 0:3] = lo.astype(np.float32)
    mesh = Mesh(ctx, pos)
    lbvh = Bvh.lbvh(ctx, mesh)
    phr = Bvh.phr(ctx, lbvh)
    olbvh = oracle.Tree(pos)
    trees = ((lbvh, olbvh), (phr, oracle.PhrTree(olbvh)))
    cam = scenes.camera_rays(scenes.Camera(4 / 3, 50.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5), (0, 1, 0)), 640, 480)
    try:
        for bvh, otree in trees:
            set_tune(ctx, trav_threads=128, trav_persistent=True, trav_order=False)
            ref_hit, ref_t = bvh.closest_hit(rays)
            set_tune(ctx, trav_threads=128, trav_persistent=True, trav_order=True)
            hit, t = bvh.closest_hit(rays)
            assert np.array_equal(hit, ref_hit) and np.array_equal(t, ref_t)
            ohit, ot = otree.closest_hit(rays)
            _check(hit, t, ohit, ot, min_hits=0.5)
            for n in (2**18 - 1, 2**18, 2**18 + 1, 2**18 + 4095):
                h, tt = bvh.closest_hit(rays[:n])
                assert np.array_equal(h, ref_hit[:n]) and np.array_equal(tt, ref_t[:n]), n
            crowded = rays.copy()
            crowded[: int(0.6 * len(rays)), 0:3] = np.float32(0.05)
            for batch in (cam, crowded):
                set_tune(ctx, trav_threads=128, trav_persistent=True, trav_order=False)
                ref_hit, ref_t = bvh.closest_hit(batch)
                set_tune(ctx, trav_threads=128, trav_persistent=True, trav_order=True)
                hit, t = bvh.closest_hit(batch)
                assert np.array_equal(hit, ref_hit) and np.array_equal(t, ref_t)
                ohit, ot = otree.closest_hit(batch)
                _check(hit, t, ohit, ot, min_hits=0.05)
    finally:
        set_tune(ctx)


def test_origin_order_degenerate_batches(ctx):
    # The origin cells of a flat scene (a root box of zero extent on one axis)
    # and of rays with NaN or infinite origins are defined (cell 0 or the
    # clamped edge cell): the ordered trace equals the input-order trace bit
    # for bit.  (Coplanar overlapping triangles tie in t almost everywhere, and
    # FP32 and FP64 break those ties differently, so this scene is not one to
    # compare with the oracle.)
    print("traverse origin order, degenerate batches", flush=True)
    pos = scenes.random_triangles(10_000, seed=45, mean_edge=0.02)
    pos[:, 2::3] = 0.5  # every vertex at z = 0.5
    rays = scenes.incoherent_rays(2**18 + 7, seed=46)
    rays[:, 2] = np.float32(0.25)
    rays[::97, 0] = np.nan
    rays[1::97, 1] = np.inf
    rays[2::97, 2] = -np.inf
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    try:
        set_tune(ctx, trav_threads=128, trav_persistent=True, trav_order=False)
        ref_hit, ref_t = bvh.closest_hit(rays)
        set_tune(ctx, trav_threads=128, trav_persistent=True, trav_order=True)
        hit, t = bvh.closest_hit(rays)
        # a non-finite origin makes every Moeller-Trumbore comparison false, so
        # such a ray can end on a triangle with a NaN distance, in both orders
        assert np.array_equal(hit, ref_hit) and np.array_equal(t, ref_t, equal_nan=True)
        assert np.mean(hit >= 0) > 0.05
    finally:
        set_tune(ctx)


def test_empty_ray_batch(ctx):
    pos = scenes.random_triangles(10, seed=1)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    hit, t = bvh.closest_hit(np.zeros((0, 8), np.float32))
    assert hit.size == 0 and t.size == 0


@pytest.mark.parametrize("key_mode", [KEY_REF63, KEY_M30])
def test_config2_sampled(ctx, oracle, key_mode):
    """1M triangles; 200k incoherent + 200k coherent rays checked ray by ray."""
    print(f"traverse 1M sampled mode={key_mode}", flush=True)
    pos = scenes.random_triangles(1_000_000, seed=1)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos), key_mode=key_mode)
    tree = oracle.Tree(pos, key_mode=key_mode)
    inc = scenes.incoherent_rays(200_000, seed=2)
    hit, t = bvh.closest_hit(inc)
    ohit, ot = tree.closest_hit(inc)
    _check(hit, t, ohit, ot, min_hits=0.8)
    coh = scenes.coherent_rays_cfg2(4096, 2048)[::41]
    hit, t = bvh.closest_hit(coh)
    ohit, ot = tree.closest_hit(coh)
    _check(hit, t, ohit, ot, min_hits=0.3)


def test_config2_full_batch_properties(ctx):
    """All 8M incoherent rays: every reported hit re-intersects its triangle at
    the reported t in FP64, and t >= tmin."""
    print("traverse 8M properties", flush=True)
    pos = scenes.random_triangles(1_000_000, seed=1)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    rays = scenes.incoherent_rays(8_000_000, seed=2)
    hit, t = bvh.closest_hit(rays)
    h = hit >= 0
    assert h.mean() > 0.8
    assert np.all(t[h] >= rays[h, 3])
    sel = np.flatnonzero(h)[::97]
    tri = pos[hit[sel]].reshape(-1, 3, 3)
    o = rays[sel, 0:3].astype(np.float64)
    d = rays[sel, 4:7].astype(np.float64)
    e1 = tri[:, 1] - tri[:, 0]
    e2 = tri[:, 2] - tri[:, 0]
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    tv = o - tri[:, 0]
    q = np.cross(tv, e1)
    tt = np.einsum("ij,ij->i", e2, q) / det
    assert np.max(np.abs(tt - t[sel]) / np.abs(tt)) < T_REL


def test_axis_parallel_rays_stay_cheap(ctx, oracle):
    """Rays with zero direction components (1/d = +-inf) must keep full slab
    culling: they visit about as many nodes as slightly tilted copies, and
    their hits match the oracle."""
    print("traverse axis-parallel rays", flush=True)
    pos = scenes.random_triangles(50_000, seed=25, mean_edge=0.01)
    rng = np.random.default_rng(26)
    n = 20_000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(0, 1, (n, 3))
    rays[:, 3] = 0.001
    axis = rng.integers(0, 3, n)
    rays[np.arange(n), 4 + axis] = rng.choice([-1.0, 1.0], n)
    rays[:, 7] = np.inf
    tilted = rays.copy()
    tilted[:, 4:7] += np.float32(1e-3) * rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    nodes, _ = bvh.counts(rays)
    tnodes, _ = bvh.counts(tilted)
    assert nodes.mean() <= 2.0 * tnodes.mean(), (nodes.mean(), tnodes.mean())
    hit, t = bvh.closest_hit(rays)
    ohit, ot = oracle.Tree(pos).closest_hit(rays)
    _check(hit, t, ohit, ot, min_hits=0.0)


def test_leaf_larger_than_wide_count_field(ctx, oracle):
    """70,000 copies of one triangle share one Morton key, so the LBVH has a
    leaf of 70,000 triangles -- beyond the 16-bit count of a wide-tree child;
    the widening pass splits it into a 4-ary sub-tree of small leaves.  Hits
    (the tie goes to the last copy) and t match the oracle."""
    print("traverse huge leaf", flush=True)
    base = scenes.random_triangles(2000, seed=27, mean_edge=0.02)
    dup = np.repeat(base[:1], 70_000, axis=0)
    pos = np.concatenate([base[:1000], dup, base[1000:]])
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    c = dup[0].reshape(3, 3).mean(0)
    rng = np.random.default_rng(28)
    n = 4000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = c + rng.normal(0, 0.3, (n, 3))
    rays[:, 4:7] = (c + rng.normal(0, 0.002, (n, 3))) - rays[:, :3]
    rays[:, 3] = 0.001
    rays[:, 7] = np.inf
    rays = np.concatenate([rays, scenes.incoherent_rays(4000, seed=29)])
    hit, t = bvh.closest_hit(rays)
    ohit, ot = oracle.Tree(pos).closest_hit(rays)
    _check(hit, t, ohit, ot, min_hits=0.3)
    assert np.any((hit >= 1000) & (hit < 71_000))  # the big leaf is hit
