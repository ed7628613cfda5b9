"""GPU parity of the render adjoint (path replay) and the per-triangle
parameters it differentiates.

Bars: the render with bound per-triangle albedo / emission matches the FP64
oracle within 3 sigma per pixel (>= 99 %); binding the material colours as
parameters changes nothing; the gradient agrees with the FP64 oracle gradient
to 1e-2 (norm-wise relative error, and per triangle on the largest entries),
with central differences of the GPU's own renders (h = 1e-3, common random
numbers) to 1e-2, and with the closed forms of a furnace sphere and a
directly viewed light."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Mesh, render, render_backward, render_device, render_segments

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    torch.cuda.synchronize()
    return t


class Scene:
    def __init__(self, ctx, d):
        self.ctx = ctx
        self.d = d
        self.mesh = Mesh(ctx, d["pos"])
        self.mesh.set_shading(d["mat_idx"], d["materials"])
        self.bvh = Bvh.lbvh(ctx, self.mesh)
        self.n = len(d["pos"])

    def bind(self, alb, emi):
        self.alb = _dev(alb)
        self.emi = _dev(emi)
        self.mesh.bind_params(self.alb.data_ptr(), self.emi.data_ptr())

    def mean(self, W, H, spp, depth, miss, seed):
        s = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        q = torch.empty_like(s)
        torch.cuda.synchronize()
        render_device(self.ctx, self.mesh, self.bvh, self.d["camera"], W, H, spp, depth, miss, seed, s.data_ptr(),
                      q.data_ptr())
        self.ctx.synchronize()
        return s.cpu().numpy().astype(np.float64) / spp

    def grad(self, W, H, spp, depth, miss, seed, g_img):
        g = _dev(g_img)
        ga = torch.empty((self.n, 3), dtype=torch.float32, device="cuda")
        ge = torch.empty_like(ga)
        torch.cuda.synchronize()
        render_backward(self.ctx, self.mesh, self.bvh, self.d["camera"], W, H, spp, depth, miss, seed, g.data_ptr(),
                        ga.data_ptr(), ge.data_ptr())
        self.ctx.synchronize()
        return ga.cpu().numpy().astype(np.float64), ge.cpu().numpy().astype(np.float64)


def _within_3_sigma(s1, q1, s2, q2, n, floor=1e-4):
    m1, m2 = s1 / n, s2 / n
    v1 = np.maximum(q1 / n - m1 * m1, 0) * n / max(n - 1, 1)
    v2 = np.maximum(q2 / n - m2 * m2, 0) * n / max(n - 1, 1)
    se = np.sqrt(v1 / n + v2 / n)
    diff = np.abs(m1 - m2)
    return float(((diff > 3 * se) & (diff > floor + 1e-5 * np.abs(m2))).mean())


W, H, SPP, DEPTH = 64, 48, 8, 3


@pytest.fixture(scope="module")
def small(ctx):
    d = scenes.adjoint_scene(4000, seed=7, width=W, height=H)
    d["camera"] = scenes.Camera(W / H, 45.0, (0.0, 0.0, 3.4), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    return Scene(ctx, d)


def test_bound_params_render_matches_oracle(ctx, oracle, small):
    print("adjoint: primal with per-triangle params vs oracle", flush=True)
    d = small.d
    small.bind(d["albedo"], d["emission"])
    s, q, _ = render(ctx, small.mesh, small.bvh, d["camera"], W, H, SPP, DEPTH, _lib.MISS_DEFAULT, 3)
    os_, oq, _ = oracle.render(oracle.Tree(d["pos"]), d["mat_idx"], d["materials"], d["camera"], W, H, SPP, DEPTH,
                               _lib.MISS_DEFAULT, 3, albedo=d["albedo"], emission=d["emission"])
    assert _within_3_sigma(s, q, os_, oq, SPP) <= 0.01
    assert s.max() > 0


def test_binding_material_colours_changes_nothing(ctx, small):
    d = small.d
    small.mesh.bind_params(None, None)
    a = render(ctx, small.mesh, small.bvh, d["camera"], W, H, 2, DEPTH, _lib.MISS_SKY, 5)
    mats = d["materials"]
    small.bind(mats[d["mat_idx"], 1:4], mats[d["mat_idx"], 1:4])
    b = render(ctx, small.mesh, small.bvh, d["camera"], W, H, 2, DEPTH, _lib.MISS_SKY, 5)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("zeros", [False, True])
def test_gradient_matches_oracle(ctx, oracle, small, zeros):
    print(f"adjoint: gradient vs FP64 oracle zeros={zeros}", flush=True)
    d = small.d
    alb = d["albedo"].copy()
    if zeros:  # zero albedo channels exercise the zero-count rule of the replay
        rng = np.random.default_rng(4)
        alb[rng.random(len(alb)) < 0.3, rng.integers(0, 3)] = 0.0
        alb[:12][::3, 1] = 0.0
    small.bind(alb, d["emission"])
    g_img = np.random.default_rng(5).normal(size=(H, W, 3))
    ga, ge = small.grad(W, H, SPP, DEPTH, _lib.MISS_SKY, 9, g_img)
    _, oga, oge = oracle.render_grad(oracle.Tree(d["pos"]), d["mat_idx"], d["materials"], d["camera"], W, H, SPP,
                                     DEPTH, _lib.MISS_SKY, 9, g_img, albedo=alb, emission=d["emission"])
    for got, want in ((ga, oga), (ge, oge)):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err <= 1e-2, err
    big = np.abs(oga) > 0.1 * np.abs(oga).max()
    assert big.sum() >= 3
    assert np.all(np.abs(ga[big] - oga[big]) <= 1e-2 * np.abs(oga[big]) + 1e-3 * np.abs(oga).max())


def test_gradient_matches_central_differences(ctx, small):
    """dLoss/da by central differences of the GPU's own renders (same seed, so
    the same paths), loss = sum(g * mean image)."""
    print("adjoint: gradient vs central differences", flush=True)
    d = small.d
    alb = d["albedo"].copy()
    emi = d["emission"].copy()
    small.bind(alb, emi)
    g_img = np.random.default_rng(6).normal(size=(H, W, 3))
    ga, ge = small.grad(W, H, SPP, DEPTH, _lib.MISS_SKY, 13, g_img)
    order = np.argsort(-np.abs(ga).max(1))
    picks = [("a", int(t), int(np.argmax(np.abs(ga[t])))) for t in order[:12]]
    light = int(np.nonzero(d["mat_idx"] == 3)[0][0])
    picks.append(("e", light, 1))
    h = 1e-3
    for which, tri, ch in picks:
        vals, steps = [], []
        for sign in (1, -1):
            a, e = alb.copy(), emi.copy()
            p = a if which == "a" else e
            p[tri, ch] = np.float32(p[tri, ch] + sign * h)
            steps.append(float(p[tri, ch]))
            small.bind(a, e)
            vals.append(float(np.sum(g_img * small.mean(W, H, SPP, DEPTH, _lib.MISS_SKY, 13))))
        fd = (vals[0] - vals[1]) / (steps[0] - steps[1])
        want = (ga if which == "a" else ge)[tri, ch]
        assert abs(fd - want) <= 1e-2 * abs(want), (which, tri, ch, fd, want)
    small.bind(alb, emi)


def test_furnace_gradient_sum(ctx):
    """Convex sphere, albedo a everywhere, white miss, loss = sum of the mean
    image: sum_t dL/da_t = (W*H - sum(mean)) / (1 - a)."""
    pos = scenes.uv_sphere((0, 0, 0), 1.0, 24, 14).astype(np.float32).astype(np.float64)
    n = len(pos)
    a = 0.6
    d = {"pos": pos, "mat_idx": np.zeros(n, np.int32),
         "materials": np.array([[scenes.MAT_DIFFUSE, 0.3, 0.3, 0.3]], np.float32),
         "camera": scenes.Camera(1.5, 60.0, (0, 0, 3.0), (0, 0, 0), (0, 1, 0))}
    sc = Scene(ctx, d)
    sc.bind(np.full((n, 3), a, np.float32), np.zeros((n, 3), np.float32))
    w, h, spp = 96, 64, 4
    mean = sc.mean(w, h, spp, 4, _lib.MISS_WHITE, 21)
    primal_segments = render_segments(ctx)
    ga, ge = sc.grad(w, h, spp, 4, _lib.MISS_WHITE, 21, np.ones((h, w, 3)))
    for c in range(3):
        assert ga[:, c].sum() == pytest.approx((w * h - mean[..., c].sum()) / (1 - a), rel=1e-4)
    assert np.all(ge == 0) and np.all(ga >= 0)
    # every path ends in the white sky and carries a gradient: the replay
    # sweep retraces each one, so the backward traces twice the primal's rays
    assert primal_segments > w * h * spp
    assert render_segments(ctx) == 2 * primal_segments


def test_direct_light_emission_gradient(ctx):
    quad = np.array([[-1, -1, 0, 1, -1, 0, 1, 1, 0], [-1, -1, 0, 1, 1, 0, -1, 1, 0]], np.float64)
    d = {"pos": quad, "mat_idx": np.zeros(2, np.int32),
         "materials": np.array([[scenes.MAT_LIGHT, 1.0, 1.0, 1.0]], np.float32),
         "camera": scenes.Camera(1.0, 60.0, (0.2, 0.1, 2.5), (0, 0, 0), (0, 1, 0))}
    sc = Scene(ctx, d)
    e = np.array([[2.0, 3.0, 5.0]] * 2, np.float32)
    sc.bind(np.zeros((2, 3), np.float32), e)
    mean = sc.mean(40, 40, 2, 2, _lib.MISS_DEFAULT, 4)
    ga, ge = sc.grad(40, 40, 2, 2, _lib.MISS_DEFAULT, 4, np.ones((40, 40, 3)))
    for c in range(3):
        assert ge[:, c].sum() == pytest.approx(mean[..., c].sum() / e[0, c], rel=1e-4)
    assert np.all(ga == 0)


def test_backward_argument_checks(ctx, small):
    g = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    out = torch.zeros((small.n, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.TraceError):
        render_backward(ctx, small.mesh, small.bvh, small.d["camera"], W, H, 0, DEPTH, 0, 1, g.data_ptr(),
                        out.data_ptr(), out.data_ptr())
    with pytest.raises(_lib.TraceError):
        render_backward(ctx, small.mesh, small.bvh, small.d["camera"], W, H, 1, DEPTH, 0, 1, 0, out.data_ptr(),
                        out.data_ptr())


def test_adam_step_matches_reference(ctx):
    """trace_adam_step against the textbook update in FP64 (Kingma & Ba), with
    clamping, over three steps."""
    from trace_amd.device import adam_step

    rng = np.random.default_rng(8)
    n = 10_001
    p0 = rng.uniform(0, 1, n).astype(np.float32)
    grads = [rng.normal(0, 1, n).astype(np.float32) for _ in range(3)]
    p, m, v = _dev(p0), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    rp, rm, rv = p0.astype(np.float64), np.zeros(n), np.zeros(n)
    lr, b1, b2, eps = 0.05, 0.9, 0.999, 1e-8
    for t, g in enumerate(grads, start=1):
        dg = _dev(g)
        adam_step(ctx, p.data_ptr(), dg.data_ptr(), m.data_ptr(), v.data_ptr(), n, lr, t, b1, b2, eps, 0.0, 1.0)
        ctx.synchronize()
        rm = b1 * rm + (1 - b1) * g
        rv = b2 * rv + (1 - b2) * g.astype(np.float64) ** 2
        rp = np.clip(rp - lr * (rm / (1 - b1 ** t)) / (np.sqrt(rv / (1 - b2 ** t)) + eps), 0.0, 1.0)
    got = p.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got - rp)) <= 1e-5
    with pytest.raises(_lib.TraceError):
        adam_step(ctx, p.data_ptr(), p.data_ptr(), m.data_ptr(), v.data_ptr(), n, lr, 0)


def test_inverse_rendering_lowers_the_loss(ctx, small):
    """Render a target from the ground-truth albedo, start from 0.5 and take
    Adam steps on dL/d(albedo) of the L2 image loss: the loss falls."""
    from trace_amd.device import adam_step

    print("adjoint: inverse rendering steps", flush=True)
    d = small.d
    small.bind(d["albedo"], d["emission"])
    target = small.mean(W, H, SPP, DEPTH, _lib.MISS_SKY, 31)
    n = small.n
    alb = torch.full((n, 3), 0.5, dtype=torch.float32, device="cuda")
    m = torch.zeros_like(alb)
    v = torch.zeros_like(alb)
    small.alb = alb
    small.mesh.bind_params(alb.data_ptr(), small.emi.data_ptr())
    losses = []
    for step in range(1, 9):
        img = small.mean(W, H, SPP, DEPTH, _lib.MISS_SKY, 31)
        losses.append(float(np.mean((img - target) ** 2)))
        ga, _ = small.grad(W, H, SPP, DEPTH, _lib.MISS_SKY, 31, 2.0 * (img - target) / img.size)
        g = _dev(ga)
        adam_step(ctx, alb.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), 3 * n, 0.02, step)
        ctx.synchronize()
    assert losses[-1] < 0.8 * losses[0], losses


@pytest.mark.parametrize("zeros,miss,batch,order", [(False, _lib.MISS_DEFAULT, None, 1),
                                                     (True, _lib.MISS_SKY, None, 1),
                                                     (True, _lib.MISS_SKY, 2 * W * H, 1),
                                                     (True, _lib.MISS_SKY, 2 * W * H, 2)])
def test_kept_summaries_give_the_same_image_and_gradient(ctx, small, zeros, miss, batch, order):
    """trace_render_keep derives the image from the backward's summary sweep:
    it equals trace_render_device's bit for bit (zero albedo channels exercise
    the zero counts; the sky makes every escaping path radiant).  The backward
    from the kept summaries traces only the replay and gives the same gradient
    as the self-contained backward, up to the order of its atomic sums.  With
    a wavefront batch of 2 W H paths the 5 samples run as batches of 2 + 2 + 1,
    so the keep buffer's sample-major offsets of the later batches are used.
    render_order 2 groups every queue after the camera's by origin cell, which
    the default leaves to batches of 2^22 paths or more: the same bar holds."""
    from trace_amd.device import set_tune

    print(f"adjoint: kept summaries zeros={zeros} miss={miss} batch={batch} order={order}", flush=True)
    if batch is not None or order != 1:
        set_tune(ctx, render_batch=batch or (1 << 27), render_order=order)
    try:
        _kept_summaries_case(ctx, small, zeros, miss)
    finally:
        set_tune(ctx)


def _kept_summaries_case(ctx, small, zeros, miss):
    from trace_amd.device import render_backward_kept, render_keep

    d = small.d
    alb = d["albedo"].copy()
    if zeros:
        rng = np.random.default_rng(9)
        alb[rng.random(len(alb)) < 0.3, rng.integers(0, 3)] = 0.0
    small.bind(alb, d["emission"])
    spp, seed = 5, 21  # 5 samples: the keep buffer's sample-major layout over an odd count
    out = {k: torch.zeros((H, W, 3), dtype=torch.float32, device="cuda") for k in ("s", "q", "l", "s2", "q2", "l2")}
    keep = torch.full((spp, H, W, 4), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    render_device(ctx, small.mesh, small.bvh, d["camera"], W, H, spp, DEPTH, miss, seed, out["s"].data_ptr(),
                  out["q"].data_ptr(), out["l"].data_ptr())
    render_keep(ctx, small.mesh, small.bvh, d["camera"], W, H, spp, DEPTH, miss, seed, out["s2"].data_ptr(),
                out["q2"].data_ptr(), keep.data_ptr(), out["l2"].data_ptr())
    ctx.synchronize()
    for a, b in (("s", "s2"), ("q", "q2"), ("l", "l2")):
        assert torch.equal(out[a], out[b]), a
    assert not torch.isnan(keep).any()  # every path's summary was written
    g_img = np.random.default_rng(3).standard_normal((H, W, 3)) * 0.1
    ga, ge = small.grad(W, H, spp, DEPTH, miss, seed, g_img)
    full_segments = render_segments(ctx)
    g = _dev(g_img)
    ka = torch.empty((small.n, 3), dtype=torch.float32, device="cuda")
    ke = torch.empty_like(ka)
    torch.cuda.synchronize()
    render_backward_kept(ctx, small.mesh, small.bvh, d["camera"], W, H, spp, DEPTH, miss, seed, keep.data_ptr(),
                         g.data_ptr(), ka.data_ptr(), ke.data_ptr())
    ctx.synchronize()
    kept_segments = render_segments(ctx)
    ka, ke = ka.cpu().numpy().astype(np.float64), ke.cpu().numpy().astype(np.float64)
    # both are float32 atomic sums over the replayed paths, queued in a different
    # order: entries where contributions cancel differ by a few float32 roundings
    # of the largest terms; a single path's share is ~1e-3 of the largest entry
    for x, y in ((ka, ga), (ke, ge)):
        assert np.abs(y).max() > 0
        assert np.allclose(x, y, rtol=1e-5, atol=1e-5 * np.abs(y).max())
    assert 0 < kept_segments < full_segments
    small.mesh.bind_params(None, None)


def test_kept_entry_points_reject_bad_input(ctx, small):
    from trace_amd.device import render_backward_kept, render_keep

    d = small.d
    s = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    keep = torch.zeros((2 * H * W * 4 + 1,), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.TraceError):  # misaligned summary buffer
        render_keep(ctx, small.mesh, small.bvh, d["camera"], W, H, 2, DEPTH, 0, 1, s.data_ptr(), s.data_ptr(),
                    keep.data_ptr() + 4)
    with pytest.raises(_lib.TraceError):
        render_backward_kept(ctx, small.mesh, small.bvh, d["camera"], W, H, 2, DEPTH, 0, 1, 0, s.data_ptr(),
                             s.data_ptr(), s.data_ptr())
    sc = scenes.specular_box_scene(slices=8, rings=6, width=16, height=16)
    mesh = Mesh(ctx, sc["pos"])
    mesh.set_shading_ex(sc["mat_idx"], sc["materials"], sc["normals"])
    bvh = Bvh.lbvh(ctx, mesh)
    with pytest.raises(_lib.TraceError):  # Reflective / Refractive are not differentiated
        render_keep(ctx, mesh, bvh, sc["camera"], 8, 8, 1, 2, 0, 1, s.data_ptr(), s.data_ptr(), keep.data_ptr())
