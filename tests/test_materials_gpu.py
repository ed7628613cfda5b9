"""GPU parity of the Reflective and Refractive materials (material.go:44-106)
against the FP64 oracle: the closed-form scenes of the oracle's own tests
(flat mirror under a half-plane light, a Ratio-1 pane, a mirror whose normals
send reflections below the surface) and per-pixel 3-sigma agreement on a box
with a mirror ball and a glass ball."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Mesh, render

pytestmark = pytest.mark.gpu

# torch is imported before any test initialises HIP through the library
torch = pytest.importorskip("torch")

from test_oracle_materials import REFLECTIVE, REFRACTIVE, _normals, _quad  # noqa: E402
from test_render_gpu import within_3_sigma  # noqa: E402


def _setup(ctx, pos, mi, mats8, nrm):
    mesh = Mesh(ctx, pos)
    mesh.set_shading_ex(mi, mats8, nrm)
    return mesh, Bvh.lbvh(ctx, mesh)


def test_flat_mirror(ctx, oracle):
    print("materials: flat mirror", flush=True)
    mirror = _quad(-4, 4, -4, 4, 0.0)
    light = _quad(-50, 0, -50, 50, 3.0, up=False)
    pos = np.concatenate([mirror, light])
    nrm = np.concatenate([_normals(mirror, (0, 0, 1)), _normals(light, (0, 0, -1))])
    mi = np.array([0, 0, 1, 1], np.int32)
    mats = np.array([[REFLECTIVE, 0.6, 0.6, 0.6, 0, 0, 0, 0], [_lib.MAT_LIGHT, 5, 4, 3, 0, 0, 0, 0]], np.float32)
    cam = scenes.Camera(1.0, 40.0, (0.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    mesh, bvh = _setup(ctx, pos, mi, mats, nrm)
    s, _, _ = render(ctx, mesh, bvh, cam, 16, 16, 4, 3, _lib.MISS_WHITE, seed=1)
    os_, _, _ = oracle.render(oracle.Tree(pos), mi, mats, cam, 16, 16, 4, 3, _lib.MISS_WHITE, seed=1, normals=nrm)
    assert np.allclose(s, os_, rtol=1e-5, atol=1e-5)
    # both halves present: albedo * emission and albedo * white, 4 samples each
    vals = s[..., 0] / 4
    assert np.isclose(vals.max(), 0.6 * 5, rtol=1e-6) and np.isclose(vals.min(), 0.6, rtol=1e-6)


def test_ratio_one_pane_and_dark_mirror(ctx, oracle):
    print("materials: pane and dark mirror", flush=True)
    pane = _quad(-3, 3, -3, 3, 0.0)
    light = _quad(-50, 50, -50, 50, -1.0)
    pos = np.concatenate([pane, light])
    nrm = np.concatenate([_normals(pane, (0, 0, 1)), _normals(light, (0, 0, 1))])
    mi = np.array([0, 0, 1, 1], np.int32)
    mats = np.array([[REFRACTIVE, 0.7, 0.7, 0.7, 1.0, 0, 0, 0], [_lib.MAT_LIGHT, 2, 3, 4, 0, 0, 0, 0]], np.float32)
    cam = scenes.Camera(1.0, 10.0, (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    mesh, bvh = _setup(ctx, pos, mi, mats, nrm)
    s, _, _ = render(ctx, mesh, bvh, cam, 8, 8, 8, 4, _lib.MISS_DEFAULT, seed=2)
    assert np.allclose(s / 8, np.array([2, 3, 4]) * np.float32(0.7), rtol=1e-6, atol=0)
    mirror = _quad(-4, 4, -4, 4, 0.0)
    mats = np.array([[REFLECTIVE, 0.9, 0.9, 0.9, 0, 0, 0, 0]], np.float32)
    cam = scenes.Camera(1.0, 30.0, (0.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    mesh, bvh = _setup(ctx, mirror, np.zeros(2, np.int32), mats, _normals(mirror, (0, 0, -1)))
    s, _, _ = render(ctx, mesh, bvh, cam, 8, 8, 2, 3, _lib.MISS_WHITE, seed=3)
    assert np.array_equal(s, np.zeros_like(s))


def test_specular_box_scene(ctx, oracle):
    print("materials: mirror and glass balls in the box", flush=True)
    sc = scenes.specular_box_scene()
    W, H, spp = sc["width"], sc["height"], 16
    mesh, bvh = _setup(ctx, sc["pos"], sc["mat_idx"], sc["materials"], sc["normals"])
    s, q, _ = render(ctx, mesh, bvh, sc["camera"], W, H, spp, 6, _lib.MISS_DEFAULT, seed=5)
    os_, oq, _ = oracle.render(oracle.Tree(sc["pos"]), sc["mat_idx"], sc["materials"], sc["camera"], W, H, spp,
                               6, _lib.MISS_DEFAULT, seed=5, normals=sc["normals"])
    frac, _ = within_3_sigma(s, q, os_, oq, spp)
    assert frac < 0.01, frac
    assert abs(s.mean() - os_.mean()) / os_.mean() < 0.01


def test_specular_materials_are_checked(ctx):
    sc = scenes.specular_box_scene(slices=8, rings=6, width=16, height=16)
    mesh = Mesh(ctx, sc["pos"])
    bad = sc["materials"].copy()
    bad[5, 4] = 0.0  # Ratio 0
    with pytest.raises(_lib.TraceError):
        mesh.set_shading_ex(sc["mat_idx"], bad, sc["normals"])
    with pytest.raises(_lib.TraceError):  # the 4-float entry point takes Diffuse and Light only
        mesh.set_shading(sc["mat_idx"], sc["materials"][:, :4], sc["normals"])
    mesh.set_shading_ex(sc["mat_idx"], sc["materials"], sc["normals"])
    bvh = Bvh.lbvh(ctx, mesh)
    from trace_amd.device import render_backward

    g = torch.zeros((16, 16, 3), dtype=torch.float32, device="cuda")
    out = torch.zeros((len(sc["pos"]), 3), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.TraceError):
        render_backward(ctx, mesh, bvh, sc["camera"], 16, 16, 1, 3, _lib.MISS_DEFAULT, 1, g.data_ptr(),
                        out.data_ptr(), out.data_ptr())
