"""Two-level (instanced) trees against the flattened scene.

The reference flattens every placement of a mesh (Scene.collectTracables,
scene.go:100-114).  The flattened scene here is built from the same FP32
object-space triangles moved by the same transforms in FP64, and the FP64
oracle answers closest hit on it.  Bar: hit IDs (flattened indices) agree on
>= 99.99 % of rays, and t within 1e-5 relative where both hit the same
triangle; rays are moved into instance space in FP64 and rounded to FP32 there,
so disagreements are confined to triangle edges."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Mesh, render, set_tune

pytestmark = pytest.mark.gpu


def _place(obj, xf):
    """Object-space triangles (n, 9) moved by rows (A | t) in FP64."""
    v = obj.reshape(-1, 3)
    A, t = xf.reshape(3, 4)[:, :3], xf.reshape(3, 4)[:, 3]
    return (v @ A.T + t).reshape(-1, 9)


def _flatten(meshes, inst_blas, xforms):
    parts, base, off = [], [], 0
    for j, xf in zip(inst_blas, xforms):
        parts.append(_place(meshes[j], xf))
        base.append(off)
        off += len(meshes[j])
    return np.concatenate(parts), np.array(base, np.int64)


def _check(hit, t, ohit, ot, min_hits):
    agree = np.mean(hit == ohit)
    assert agree >= 0.9999, f"hit agreement {agree:.6f}"
    both = (hit == ohit) & (hit >= 0)
    assert both.sum() >= min_hits
    assert np.max(np.abs(t[both] - ot[both]) / ot[both]) < 1e-5


def _rays_in(box_lo, box_hi, n, seed):
    r = scenes.incoherent_rays(n, seed=seed)
    r[:, :3] = box_lo + r[:, :3] * (box_hi - box_lo)
    return r


def test_instances_terrain_and_spheres(ctx, oracle):
    """A 256 x 256 terrain (one instance, identity) and 300 scale-and-translate
    placements of one 400-triangle sphere, as the large configuration is made."""
    print("instances: terrain + 300 spheres", flush=True)
    d = scenes.terrain_scene(n=256, n_spheres=300, width=64, height=36)
    meshes, _, _, node_mesh, mats = scenes.terrain_scene_graph(n=256, n_spheres=300)
    terrain = meshes[0].astype(np.float32).astype(np.float64)
    unit = meshes[1].astype(np.float32).astype(np.float64)
    xforms = mats.reshape(-1, 4, 4)[:, :3, :].reshape(-1, 12)
    flat, base = _flatten([terrain, unit], node_mesh, xforms)
    blas = [Bvh.lbvh(ctx, Mesh(ctx, terrain)), Bvh.lbvh(ctx, Mesh(ctx, unit))]
    inst = Bvh.instances(ctx, blas, node_mesh, xforms, base)
    assert inst.info().n_prims == len(node_mesh)
    lo, hi = flat.reshape(-1, 3).min(0), flat.reshape(-1, 3).max(0)
    rays = np.concatenate([_rays_in(lo, hi, 60_000, 71),
                           scenes.camera_rays(d["camera"], 192, 108).astype(np.float32)])
    hit, t = inst.closest_hit(rays)
    otree = oracle.Tree(flat)
    ohit, ot = otree.closest_hit(rays)
    _check(hit, t, ohit, ot, 10_000)
    assert np.any(hit >= len(terrain))  # spheres are hit
    # the flattened scene's own tree on the GPU hits the same triangles (its t
    # differs more: it rounds the small spheres' world-space vertices to FP32)
    fhit, _ = Bvh.lbvh(ctx, Mesh(ctx, flat)).closest_hit(rays)
    assert np.mean(hit == fhit) >= 0.9999
    # memory: the placements hold one sphere tree instead of 300 copies
    flat_bytes = Bvh.lbvh(ctx, Mesh(ctx, flat)).device_bytes()
    inst_bytes = inst.device_bytes() + sum(b.device_bytes() for b in blas)
    assert inst_bytes < flat_bytes


def test_instances_general_affine(ctx, oracle):
    """Two meshes (one of them refined by PHR) placed 60 times with rotations,
    non-uniform scales and shears; a forest's segment view as a third tree."""
    print("instances: general affine transforms", flush=True)
    rng = np.random.default_rng(72)
    m0 = scenes.random_triangles(3000, seed=73, mean_edge=0.02)
    m1 = scenes.displaced_sphere(2000, seed=74, radius=0.5).astype(np.float32).astype(np.float64)
    m2 = scenes.random_triangles(500, seed=75, mean_edge=0.05)
    forest = Bvh.lbvh(ctx, Mesh.segmented(ctx, [m1, m2]))
    blas = [Bvh.lbvh(ctx, Mesh(ctx, m0)), Bvh.phr(ctx, Bvh.lbvh(ctx, Mesh(ctx, m1)), branching=4),
            forest.segment(1)]
    meshes = [m0, m1, m2]
    n = 60
    inst_blas = rng.integers(0, 3, n).astype(np.int32)
    xforms = np.zeros((n, 12))
    for i in range(n):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        A = q @ np.diag(rng.uniform(0.3, 2.0, 3)) @ (np.eye(3) + 0.2 * np.triu(rng.normal(size=(3, 3)), 1))
        xforms[i] = np.concatenate([A, rng.uniform(-4, 4, (3, 1))], axis=1).reshape(-1)
    flat, base = _flatten(meshes, inst_blas, xforms)
    inst = Bvh.instances(ctx, blas, inst_blas, xforms, base)
    lo, hi = flat.reshape(-1, 3).min(0), flat.reshape(-1, 3).max(0)
    rays = _rays_in(lo, hi, 100_000, 76)
    hit, t = inst.closest_hit(rays)
    ohit, ot = oracle.Tree(flat).closest_hit(rays)
    _check(hit, t, ohit, ot, 10_000)
    for persistent in (False, True):  # the traversal form does not matter for instanced trees
        set_tune(ctx, trav_persistent=persistent)
        h2, t2 = inst.closest_hit(rays)
        assert np.array_equal(h2, hit) and np.array_equal(t2, t)
    set_tune(ctx)


def test_instances_edge_cases(ctx, oracle):
    """One instance (the top level is a single leaf), instances with equal
    boxes, and rays that miss everything."""
    print("instances: edge cases", flush=True)
    m = scenes.random_triangles(200, seed=77, mean_edge=0.05)
    blas = [Bvh.lbvh(ctx, Mesh(ctx, m))]
    eye = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float64)
    one = Bvh.instances(ctx, blas, [0], eye, [0])
    rays = _rays_in(np.zeros(3), np.ones(3), 5000, 78)
    hit, t = one.closest_hit(rays)
    ohit, ot = oracle.Tree(m).closest_hit(rays)
    _check(hit, t, ohit, ot, 100)
    # the same placement twice: equal boxes and equal t; the larger index wins
    two = Bvh.instances(ctx, blas, [0, 0], np.concatenate([eye, eye]), [0, 200])
    h2, t2 = two.closest_hit(rays)
    assert np.array_equal(h2[hit >= 0], hit[hit >= 0] + 200) and np.array_equal(t2, t)
    far = rays.copy()
    far[:, :3] += 100.0
    far[:, 4:7] = np.abs(far[:, 4:7])  # away from the mesh
    h3, t3 = two.closest_hit(far)
    assert np.all(h3 == -1) and np.all(np.isinf(t3))


def test_instances_refusals(ctx):
    m = scenes.random_triangles(100, seed=79)
    mesh = Mesh(ctx, m)
    blas = [Bvh.lbvh(ctx, mesh)]
    eye = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float64)
    with pytest.raises(_lib.TraceError):  # singular transform
        Bvh.instances(ctx, blas, [0], np.zeros((1, 12)), [0])
    with pytest.raises(_lib.TraceError):  # no such tree
        Bvh.instances(ctx, blas, [1], eye, [0])
    with pytest.raises(_lib.TraceError):  # spheres are not instanced
        Bvh.instances(ctx, [Bvh.lbvh(ctx, Mesh(ctx, scenes.mixed_scene(200, 20, seed=80)))], [0], eye, [0])
    inst = Bvh.instances(ctx, blas, [0], eye, [0])
    with pytest.raises(_lib.TraceError):
        Bvh.phr(ctx, inst)
    with pytest.raises(_lib.TraceError):
        inst.traversal_steps(scenes.incoherent_rays(10).astype(np.float64))
    with pytest.raises(_lib.TraceError):
        inst.counts(scenes.incoherent_rays(10))
    mesh.set_shading(np.zeros(100, np.int32), np.array([[0, 0.5, 0.5, 0.5]], np.float32))
    with pytest.raises(_lib.TraceError):
        render(ctx, mesh, inst, scenes.camera_cfg2(), 8, 8, 1, 1, 0, seed=1)


def test_instances_large_configuration(ctx, oracle):
    """The large configuration instanced: the 2048 x 2048 terrain and 5,000
    placements of one 400-triangle sphere; 200k rays (incoherent ones in the
    scene box and the camera's) against the oracle on the flattened scene."""
    print("instances: large configuration (terrain + 5000 spheres)", flush=True)
    meshes, _, _, node_mesh, mats = scenes.terrain_scene_graph()
    terrain = meshes[0].astype(np.float32).astype(np.float64)
    unit = meshes[1].astype(np.float32).astype(np.float64)
    xforms = mats.reshape(-1, 4, 4)[:, :3, :].reshape(-1, 12)
    flat, base = _flatten([terrain, unit], node_mesh, xforms)
    assert len(flat) == 10_380_418
    blas = [Bvh.lbvh(ctx, Mesh(ctx, terrain)), Bvh.lbvh(ctx, Mesh(ctx, unit))]
    inst = Bvh.instances(ctx, blas, node_mesh, xforms, base)
    d = scenes.terrain_scene(width=1920, height=1080)
    lo, hi = flat.reshape(-1, 3).min(0), flat.reshape(-1, 3).max(0)
    cam = scenes.camera_rays(d["camera"], 1920, 1080).astype(np.float32)[::21][:100_000]
    rays = np.concatenate([_rays_in(lo, hi, 100_000, 81), cam])
    hit, t = inst.closest_hit(rays)
    ohit, ot = oracle.Tree(flat).closest_hit(rays)
    _check(hit, t, ohit, ot, 60_000)
    sph = hit >= len(terrain)
    assert sph.sum() > 100  # the spheres are hit
