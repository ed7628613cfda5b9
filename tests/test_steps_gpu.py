"""GPU parity of BVH.traversalSteps (bvh.go:21-44) and the heat-map renderer
(render.go:167-197, 250-258) against the FP64 oracle: every count and every
colour equal, on the LBVH and on the PHR tree."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Mesh

pytestmark = pytest.mark.gpu


def _rays64(n, seed):
    return scenes.incoherent_rays(n, seed=seed).astype(np.float64)


@pytest.mark.parametrize("phr", [False, True])
def test_steps_match_oracle(ctx, oracle, phr):
    print(f"steps random phr={phr}", flush=True)
    pos = scenes.random_triangles(50_000, seed=51, mean_edge=0.01)
    lb = Bvh.lbvh(ctx, Mesh(ctx, pos))
    ot = oracle.Tree(pos)
    bvh, ref = (Bvh.phr(ctx, lb), oracle.PhrTree(ot)) if phr else (lb, ot)
    rays = _rays64(20_000, 52)
    got = bvh.traversal_steps(rays)
    want = ref.traversal_steps(rays)
    assert np.array_equal(got, want)
    assert want.max() > 20


@pytest.mark.parametrize("phr", [False, True])
def test_heatmap_matches_oracle(ctx, oracle, phr):
    print(f"heatmap box scene phr={phr}", flush=True)
    sc = scenes.box_scene(seed=1, slices=40, rings=22, width=160, height=90)
    lb = Bvh.lbvh(ctx, Mesh(ctx, sc["pos"]))
    ot = oracle.Tree(sc["pos"])
    bvh, ref = (Bvh.phr(ctx, lb), oracle.PhrTree(ot)) if phr else (lb, ot)
    rgb, counts = bvh.heatmap(sc["camera"], 160, 90, 20)
    orgb, ocounts = ref.heatmap(sc["camera"], 160, 90, 20)
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(rgb, orgb)
    assert counts.max() > 20 and counts.min() >= 0


def test_heatmap_terrain_phr(ctx, oracle):
    """Big PHR leaves (heightfield): the walk tests every triangle of a leaf."""
    print("heatmap terrain", flush=True)
    d = scenes.terrain_scene(n=256, n_spheres=100, width=128, height=72)
    lb = Bvh.lbvh(ctx, Mesh(ctx, d["pos"]))
    phr = Bvh.phr(ctx, lb)
    ref = oracle.PhrTree(oracle.Tree(d["pos"]))
    rgb, counts = phr.heatmap(d["camera"], 128, 72, 40)
    orgb, ocounts = ref.heatmap(d["camera"], 128, 72, 40)
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(rgb, orgb)


def test_steps_on_forest_segment(ctx, oracle):
    print("steps forest segment", flush=True)
    meshes = [scenes.displaced_sphere(n, seed=60 + k) for k, n in enumerate((3000, 700, 12_000))]
    forest = Bvh.lbvh(ctx, Mesh.segmented(ctx, meshes))
    rays = np.zeros((2000, 8), np.float64)
    rng = np.random.default_rng(61)
    rays[:, :3] = rng.normal(0, 2.0, (2000, 3))
    rays[:, 4:7] = -rays[:, :3] + rng.normal(0, 0.3, (2000, 3))
    rays[:, 3] = 0.001
    rays[:, 7] = np.inf
    for s, m in enumerate(meshes):
        got = forest.segment(s).traversal_steps(rays)
        assert np.array_equal(got, oracle.Tree(m).traversal_steps(rays)), s


def test_heatmap_rejects_bad_arguments(ctx):
    pos = scenes.random_triangles(100, seed=62)
    bvh = Bvh.lbvh(ctx, Mesh(ctx, pos))
    cam = scenes.Camera(1.0, 60.0, (0.5, 0.5, 3.0), (0.5, 0.5, 0.5), (0, 1, 0))
    with pytest.raises(_lib.TraceError):
        bvh.heatmap(cam, 1, 10, 5)
    with pytest.raises(_lib.TraceError):
        bvh.heatmap(cam, 10, 10, 0)
