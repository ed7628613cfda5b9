"""Pins of the FP64 oracle's render gradient (CPU only).

The albedo never steers a path (the scattered direction depends on the normal
and the random numbers only), so with common random numbers the mean image is
a polynomial in each albedo and emission: central differences of the oracle's
own primal render agree with the analytic gradient to ~1e-6.  A furnace sphere
and a directly viewed light give closed forms for the gradient sums."""
import numpy as np
import pytest

from trace_amd import _lib, scenes


def _scene(seed=3):
    sc = scenes.box_scene(seed=seed, slices=12, rings=8, width=32, height=24)
    n = len(sc["pos"])
    rng = np.random.default_rng(7)
    alb = rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32)
    emi = np.zeros((n, 3), np.float32)
    light = sc["mat_idx"] == 3
    emi[light] = np.float32(15.0)
    return sc, alb, emi


def test_tape_mean_equals_recursive_render(oracle):
    sc, alb, emi = _scene()
    tree = oracle.Tree(sc["pos"])
    args = (sc["mat_idx"], sc["materials"], sc["camera"], 32, 24, 4, 3, _lib.MISS_SKY, 11)
    s, _, _ = oracle.render(tree, *args, albedo=alb, emission=emi)
    mean, ga, ge = oracle.render_grad(tree, *args, np.ones((24, 32, 3)), albedo=alb, emission=emi)
    assert np.array_equal(mean, s / 4)
    assert np.all(ga >= 0) and np.all(ge >= 0)


def test_per_triangle_params_default_to_material_colours(oracle):
    sc, _, _ = _scene()
    tree = oracle.Tree(sc["pos"])
    args = (sc["mat_idx"], sc["materials"], sc["camera"], 32, 24, 2, 3, _lib.MISS_SKY, 5)
    mats = sc["materials"]
    alb = mats[sc["mat_idx"], 1:4].copy()
    emi = mats[sc["mat_idx"], 1:4].copy()
    a, _, _ = oracle.render(tree, *args)
    b, _, _ = oracle.render(tree, *args, albedo=alb, emission=emi)
    assert np.array_equal(a, b)


def _fd(oracle, tree, sc, alb, emi, g, which, tri, ch, h=1e-3):
    """Central difference of sum(g * mean image) in parameter (which, tri, ch)."""
    vals = []
    steps = []
    for sign in (1, -1):
        a, e = alb.copy(), emi.copy()
        p = a if which == "albedo" else e
        p[tri, ch] = np.float32(p[tri, ch] + sign * h)
        steps.append(float(p[tri, ch]))
        s, _, _ = oracle.render(tree, sc["mat_idx"], sc["materials"], sc["camera"], 32, 24, 4, 3, _lib.MISS_SKY,
                                11, albedo=a, emission=e, threads=8)
        vals.append(float(np.sum(g * (s / 4))))
    return (vals[0] - vals[1]) / (steps[0] - steps[1])


def test_gradient_matches_central_differences(oracle):
    sc, alb, emi = _scene()
    tree = oracle.Tree(sc["pos"])
    g = np.random.default_rng(2).normal(size=(24, 32, 3))
    _, ga, ge = oracle.render_grad(tree, sc["mat_idx"], sc["materials"], sc["camera"], 32, 24, 4, 3,
                                   _lib.MISS_SKY, 11, g, albedo=alb, emission=emi)
    # the triangles with the largest albedo gradients (walls, sphere) and the light
    order = np.argsort(-np.abs(ga).max(1))
    picks = [("albedo", int(t), c) for t in order[:6] for c in (0, 2)]
    light = int(np.nonzero(sc["mat_idx"] == 3)[0][0])
    picks += [("emission", light, 0), ("emission", light, 1)]
    for which, tri, ch in picks:
        want = (ga if which == "albedo" else ge)[tri, ch]
        got = _fd(oracle, tree, sc, alb, emi, g, which, tri, ch)
        assert want != 0
        assert abs(got - want) <= 1e-6 * abs(want) + 1e-12, (which, tri, ch, got, want)


def test_furnace_gradient_sum(oracle):
    """Convex sphere, one albedo a, white miss: a path is a (hit, escape) or 1
    (miss), so sum_t dL/da_t = #hit paths / spp = (W*H - sum(mean)) / (1 - a)
    for L = sum of the mean image."""
    pos = scenes.uv_sphere((0, 0, 0), 1.0, 16, 10).astype(np.float32).astype(np.float64)
    n = len(pos)
    a = np.float32(0.6)
    mats = np.array([[scenes.MAT_DIFFUSE, a, a, a]], np.float32)
    cam = scenes.Camera(1.5, 60.0, (0, 0, 3.0), (0, 0, 0), (0, 1, 0))
    W, H, spp = 36, 24, 3
    mean, ga, ge = oracle.render_grad(oracle.Tree(pos), np.zeros(n, np.int32), mats, cam, W, H, spp, 4,
                                      _lib.MISS_WHITE, 9, np.ones((H, W, 3)))
    for c in range(3):
        assert ga[:, c].sum() == pytest.approx((W * H - mean[..., c].sum()) / (1 - float(a)), rel=1e-12)
    assert np.all(ge == 0)
    assert (ga[:, 0] > 0).sum() > n // 4


def test_direct_light_emission_gradient(oracle):
    """A light quad seen directly: every camera path that hits it has value E,
    so d(sum of mean)/dE over the quad = #light paths / spp = sum(mean) / E."""
    quad = np.array([[-1, -1, 0, 1, -1, 0, 1, 1, 0], [-1, -1, 0, 1, 1, 0, -1, 1, 0]], np.float64)
    mats = np.array([[scenes.MAT_LIGHT, 2.0, 3.0, 5.0]], np.float32)
    cam = scenes.Camera(1.0, 60.0, (0.2, 0.1, 2.5), (0, 0, 0), (0, 1, 0))
    W, H, spp = 20, 20, 2
    mean, ga, ge = oracle.render_grad(oracle.Tree(quad), np.zeros(2, np.int32), mats, cam, W, H, spp, 2,
                                      _lib.MISS_DEFAULT, 4, np.ones((H, W, 3)))
    for c, e in enumerate((2.0, 3.0, 5.0)):
        assert ge[:, c].sum() == pytest.approx(mean[..., c].sum() / e, rel=1e-12)
    assert np.all(ga == 0) and np.all(ge[:, 0] > 0)
