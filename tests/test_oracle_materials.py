"""Pins the oracle's Reflective and Refractive materials (material.go:44-106)
with closed forms, before any GPU comparison:

- a flat mirror (Diffusion 0) at z = 0 seen straight down from z = 2, under
  a light covering the half-plane x < 0 at z = 3 (facing down): a camera ray
  with slope s along x meets the mirror at x = 2s and, reflected, reaches
  z = 3 at x = 5s, so the left half of the image is albedo * emission and the
  right half albedo * (white sky);
- a pane with Ratio 1 seen head-on passes every ray straight through
  (refract with a unit normal and ratio 1 is the identity; Schlick's
  reflectance at normal incidence is 0), so the image is albedo * emission
  of the light behind it;
- a mirror ray that leaves below the surface ends the path: value 0."""
import numpy as np

from trace_amd import _lib, scenes

REFLECTIVE, REFRACTIVE = 2, 3


def _quad(x0, x1, y0, y1, z, up=True):
    a, b, c, d = (x0, y0, z), (x1, y0, z), (x1, y1, z), (x0, y1, z)
    t = [[*a, *b, *c], [*a, *c, *d]] if up else [[*a, *c, *b], [*a, *d, *c]]
    return np.array(t, np.float64)


def _normals(tris, n):
    return np.tile(np.asarray(n, np.float64), (len(tris), 3))


def test_flat_mirror_reflects_into_the_half_plane_light(oracle):
    mirror = _quad(-4, 4, -4, 4, 0.0)
    light = _quad(-50, 0, -50, 50, 3.0, up=False)
    pos = np.concatenate([mirror, light])
    nrm = np.concatenate([_normals(mirror, (0, 0, 1)), _normals(light, (0, 0, -1))])
    mi = np.array([0, 0, 1, 1], np.int32)
    a = np.float32(0.6)
    mats = np.array([[REFLECTIVE, a, a, a, 0, 0, 0, 0], [_lib.MAT_LIGHT, 5, 4, 3, 0, 0, 0, 0]], np.float32)
    W = H = 16
    cam = scenes.Camera(1.0, 40.0, (0.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    s, _, _ = oracle.render(oracle.Tree(pos), mi, mats, cam, W, H, 4, 3, _lib.MISS_WHITE, seed=1, normals=nrm)
    img = s / 4
    # x of each pixel column's camera ray (the horizontal axis may point either way)
    xs = np.array([cam.directions(np.array([(i + 0.5) / (W - 1)]), np.array([0.5]))[0, 0] for i in range(W)])
    left = xs < -0.05
    right = xs > 0.05
    e = np.float32([5, 4, 3]).astype(np.float64) * float(a)
    assert np.allclose(img[:, left], e, rtol=0, atol=1e-12)
    assert np.allclose(img[:, right], float(a), rtol=0, atol=1e-12)


def test_ratio_one_pane_is_transparent(oracle):
    pane = _quad(-3, 3, -3, 3, 0.0)
    light = _quad(-50, 50, -50, 50, -1.0)
    pos = np.concatenate([pane, light])
    nrm = np.concatenate([_normals(pane, (0, 0, 1)), _normals(light, (0, 0, 1))])
    mi = np.array([0, 0, 1, 1], np.int32)
    a = np.float32(0.7)
    mats = np.array([[REFRACTIVE, a, a, a, 1.0, 0, 0, 0], [_lib.MAT_LIGHT, 2, 3, 4, 0, 0, 0, 0]], np.float32)
    cam = scenes.Camera(1.0, 10.0, (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    s, _, _ = oracle.render(oracle.Tree(pos), mi, mats, cam, 8, 8, 8, 4, _lib.MISS_DEFAULT, seed=2, normals=nrm)
    assert np.allclose(s / 8, np.array([2, 3, 4]) * float(a), rtol=1e-12, atol=0)


def test_mirror_ray_below_the_surface_ends_the_path(oracle):
    # seen from above the triangles are front faces (as in the first test),
    # but their vertex normals point down: the reflection of a downward ray
    # about n = (0, 0, -1) leaves with reflected . n < 0, so the path ends dark
    mirror = _quad(-4, 4, -4, 4, 0.0)
    nrm = _normals(mirror, (0, 0, -1))
    mi = np.zeros(2, np.int32)
    mats = np.array([[REFLECTIVE, 0.9, 0.9, 0.9, 0, 0, 0, 0]], np.float32)
    cam = scenes.Camera(1.0, 30.0, (0.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    s, _, _ = oracle.render(oracle.Tree(mirror), mi, mats, cam, 8, 8, 2, 3, _lib.MISS_WHITE, seed=3, normals=nrm)
    assert np.array_equal(s, np.zeros_like(s))
    # the same mirror with upward normals reflects into the white sky
    s, _, _ = oracle.render(oracle.Tree(mirror), mi, mats, cam, 8, 8, 2, 3, _lib.MISS_WHITE, seed=3,
                            normals=-nrm)
    assert np.allclose(s / 2, np.float32(0.9), rtol=0, atol=1e-12)
