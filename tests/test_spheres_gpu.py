"""GPU parity of analytic Spheres (primitive.go:25-86) mixed with triangles:
LBVH keys and topology, closest hit on the LBVH and the PHR tree, the
traversalSteps count (equal except where a ray grazes an outward-rounded
sphere box), a furnace of one sphere, a box with mirror and glass
spheres (3 sigma per pixel), scene compile of a sphere mesh, animation keeping
radii, and rejection of malformed spheres."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Mesh, render

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from test_render_gpu import within_3_sigma  # noqa: E402


def _check_hits(hit, t, ohit, ot):
    assert np.mean(hit == ohit) >= 0.9999
    both = (hit == ohit) & (hit >= 0)
    assert both.sum() > 0
    assert np.max(np.abs(t[both] - ot[both]) / ot[both]) < 1e-5


@pytest.mark.parametrize("phr", [False, True])
def test_mixed_closest_hit(ctx, oracle, phr):
    print(f"spheres: closest hit phr={phr}", flush=True)
    pos = scenes.mixed_scene(30_000, 2000, seed=84)
    lb = Bvh.lbvh(ctx, Mesh(ctx, pos))
    ot = oracle.Tree(pos)
    keys, perm = lb.sorted()
    okeys, operm = ot.sorted()
    assert np.array_equal(keys, okeys) and np.array_equal(perm, operm)
    _, box = lb.canonical()
    _, obox = ot.canonical()
    # sphere corners c -/+ r are not FP32 values: the device keeps them rounded outward
    lo32 = obox[:, :3].astype(np.float32)
    lo32 = np.where(lo32.astype(np.float64) > obox[:, :3], np.nextafter(lo32, np.float32(-np.inf)), lo32)
    hi32 = obox[:, 3:].astype(np.float32)
    hi32 = np.where(hi32.astype(np.float64) < obox[:, 3:], np.nextafter(hi32, np.float32(np.inf)), hi32)
    assert np.array_equal(box, np.concatenate([lo32, hi32], axis=1))
    bvh = Bvh.phr(ctx, lb) if phr else lb
    ref = oracle.PhrTree(ot) if phr else ot
    rays = scenes.incoherent_rays(50_000, seed=85)
    hit, t = bvh.closest_hit(rays)
    ohit, ot_ = ref.closest_hit(rays)
    _check_hits(hit, t, ohit, ot_)
    assert np.isnan(pos[hit[hit >= 0], 4]).sum() > 100
    # traversalSteps: sphere boxes (and the inner boxes holding them) are FP32
    # rounded outward on the device, so a ray grazing one may enter a box the
    # FP64 walk misses; elsewhere the counts are equal
    steps = bvh.traversal_steps(rays.astype(np.float64)[:5000])
    osteps = ref.traversal_steps(rays.astype(np.float64)[:5000])
    same = np.mean(steps == osteps)
    print(f"traversalSteps equal on {same:.4f} of rays", flush=True)
    assert same >= 0.999
    assert np.max(np.abs(steps - osteps)) <= 8


def test_sphere_furnace(ctx):
    print("spheres: furnace", flush=True)
    pos = scenes.sphere_prims([(0.0, 0.0, 0.0)], 1.0)
    mesh = Mesh(ctx, pos)
    a = np.float32(0.6)
    mesh.set_shading(np.zeros(1, np.int32), np.array([[scenes.MAT_DIFFUSE, a, a, a]], np.float32))
    cam = scenes.Camera(2.0, 60.0, (0, 0, 3.0), (0, 0, 0), (0, 1, 0))
    _, _, last = render(ctx, mesh, Bvh.lbvh(ctx, mesh), cam, 96, 48, 2, 5, _lib.MISS_WHITE, seed=7)
    vals = set(np.unique(last).tolist())
    assert vals <= {float(a), 1.0} and float(a) in vals


def test_analytic_specular_box(ctx, oracle):
    """The specular box with its two balls as analytic spheres."""
    print("spheres: mirror and glass analytic balls", flush=True)
    sc = scenes.specular_box_scene(slices=4, rings=3)
    walls = sc["pos"][:12]
    balls = scenes.sphere_prims([(-0.45, -0.55, -0.1), (0.45, -0.55, 0.25)], 0.42)
    pos = np.concatenate([walls, balls])
    nrm = np.concatenate([sc["normals"][:12], np.zeros((2, 9))])
    mi = np.concatenate([sc["mat_idx"][:12], [4, 5]]).astype(np.int32)
    mesh = Mesh(ctx, pos)
    mesh.set_shading_ex(mi, sc["materials"], nrm)
    W, H, spp = sc["width"], sc["height"], 16
    s, q, _ = render(ctx, mesh, Bvh.lbvh(ctx, mesh), sc["camera"], W, H, spp, 6, _lib.MISS_DEFAULT, seed=5)
    os_, oq, _ = oracle.render(oracle.Tree(pos), mi, sc["materials"], sc["camera"], W, H, spp, 6,
                               _lib.MISS_DEFAULT, seed=5, normals=nrm)
    frac, _ = within_3_sigma(s, q, os_, oq, spp)
    assert frac < 0.01, frac
    assert abs(s.mean() - os_.mean()) / os_.mean() < 0.01


def test_scene_compile_and_animation_keep_spheres(ctx):
    from oracle.pyoracle import scene_flatten
    from trace_amd.device import Scene

    print("spheres: scene compile and animation", flush=True)
    ball = scenes.sphere_prims([(0.1, 0.2, 0.3)], 0.25)
    tri = scenes.random_triangles(50, seed=86)
    m = np.eye(4)
    m[:3, :3] *= 2.0
    m[:3, 3] = (1.0, -1.0, 0.5)
    mats = np.stack([m.reshape(-1), np.eye(4).reshape(-1)])
    materials = np.array([[scenes.MAT_DIFFUSE, 0.5, 0.5, 0.5]], np.float32)
    sc = Scene(ctx, [ball, tri], [0, 0], materials)
    mesh = sc.compile([-1, 0], [0, 1], mats)
    opos, _, _ = scene_flatten([ball, tri], [0, 0], [-1, 0], [0, 1], mats)
    got = mesh.export()
    assert np.isnan(got[0, 4]) and got[0, 3] == 0.25  # the radius is not scaled (primitive.go:42-50)
    assert np.array_equal(np.nan_to_num(got), np.nan_to_num(opos))
    # animation moves a sphere's centre and keeps its radius and tag
    rest = Mesh(ctx, np.concatenate([ball, tri]))
    anim = Mesh(ctx, np.concatenate([ball, tri]))
    frame = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    anim.animate(rest, 0.05, 10.0, 3, frame.data_ptr())
    a = anim.export()
    assert np.isnan(a[0, 4]) and a[0, 3] == 0.25 and np.all(a[0, 5:] == 0)


def test_malformed_spheres_are_rejected(ctx):
    bad = scenes.sphere_prims([(0, 0, 0)], -1.0)
    with pytest.raises(_lib.TraceError):
        Mesh(ctx, bad)
    bad = scenes.sphere_prims([(np.inf, 0, 0)], 1.0)
    with pytest.raises(_lib.TraceError):
        Mesh(ctx, bad)
