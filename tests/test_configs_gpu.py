"""GPU parity at the benchmark workloads' own sizes.

- Batched builds and renders: the 64-mesh forest (triangle counts log-uniform
  in [1k, 500k], 4.32M triangles): the smallest, the largest and six seeded
  trees bit-exact against the FP64 oracle's LBVH of that mesh alone; scenes of
  the 32-scene batched render (512x512, 8 spp, depth 4) within 3 sigma of the
  oracle's render of their mesh alone; frame 120 of the 200k-triangle
  animation replayed from one captured HIP graph bit-exact against the
  oracle's LBVH of the exported positions and equal to an eager frame 120.
- Differentiation: the 200k-triangle scene at 512x512, 32 spp, depth 5, from
  the uniform 0.5 albedo with the L2 image loss against the ground-truth
  render: the adjoint gradient within 1e-2 (norm-wise) of the FP64 oracle's,
  and within 1e-2 of central differences (h = 1e-3, the same paths) on 256
  sampled triangles and on the light's emission.
- Large: the 10.38M-triangle terrain: LBVH and PHR trees bit-exact against the
  oracle, 200k sampled closest hits (the benchmark's persistent 128-thread
  traversal) against the oracle, the full 1920x1080x64-spp render
  repeatable bit for bit, and the scene at 96x54 within 3 sigma of the oracle.
- The 1M-triangle PHR tree with the benchmark's traversal variant, on 200k
  incoherent and 1.2M coherent rays; the Sun miss shader against the oracle."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Graph, Mesh, render, render_backward, render_batched, render_device, set_tune

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _assert_same_tree(view, tree, boxes_exact=True):
    keys, perm = view.sorted()
    okeys, operm = tree.sorted()
    assert np.array_equal(keys, okeys), "Morton keys differ"
    assert np.array_equal(perm, operm), "sort permutation differs"
    rec, box = view.canonical()
    orec, obox = tree.canonical()
    assert rec.shape == orec.shape, (rec.shape, orec.shape)
    assert np.array_equal(rec, orec), "topology differs"
    if boxes_exact:
        assert np.array_equal(box.astype(np.float64), obox), "boxes differ"
    else:  # FP32 boxes: the FP64 boxes rounded outwards
        lo, hi = obox[:, :3], obox[:, 3:]
        flo, fhi = lo.astype(np.float32), hi.astype(np.float32)
        flo = np.where(flo.astype(np.float64) > lo, np.nextafter(flo, np.float32(-np.inf)), flo)
        fhi = np.where(fhi.astype(np.float64) < hi, np.nextafter(fhi, np.float32(np.inf)), fhi)
        assert np.array_equal(box, np.concatenate([flo, fhi], axis=1)), "boxes differ"
    # Cost() is taken over the tree's FP32 boxes: equal to the FP64 one where the
    # boxes are FP32-exact, within the outward rounding (a few 1e-7 of each
    # surface) where they are not
    assert view.cost() == pytest.approx(tree.cost(), rel=1e-12 if boxes_exact else 1e-5)


def _within_3_sigma(s1, q1, s2, q2, n, floor=1e-4):
    m1, m2 = s1 / n, s2 / n
    v1 = np.maximum(q1 / n - m1 * m1, 0) * n / max(n - 1, 1)
    v2 = np.maximum(q2 / n - m2 * m2, 0) * n / max(n - 1, 1)
    se = np.sqrt(v1 / n + v2 / n)
    diff = np.abs(m1 - m2)
    return float(((diff > 3 * se) & (diff > floor + 1e-5 * np.abs(m2))).mean())


# ---- batched builds and renders (64 meshes, 32 scenes, 120 frames) ----------------------


@pytest.fixture(scope="module")
def forest64(ctx):
    meshes = scenes.batch_meshes(64, 1000, 500_000, seed=11)
    seg = Mesh.segmented(ctx, meshes)
    forest = Bvh.lbvh(ctx, seg)
    yield meshes, seg, forest
    forest.close()
    seg.close()


def test_batched_forest_trees_match_oracle(ctx, oracle, forest64):
    meshes, _, forest = forest64
    sizes = np.array([len(m) for m in meshes])
    assert sizes.sum() == 4_322_168
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    pick |= set(np.random.default_rng(3).choice(64, 6, replace=False).tolist())
    for i in sorted(pick):
        print(f"config3 forest tree {i}: {sizes[i]} tris", flush=True)
        v = forest.segment(i)
        _assert_same_tree(v, oracle.Tree(meshes[i]))
        v.close()


def test_batched_render_matches_oracle(ctx, oracle, forest64):
    """The benchmark's batched render (32 scenes, 512x512, 8 spp, depth 4, sky
    miss, scene c seeded 3 + c); the scenes of the smallest and the largest
    mesh against the oracle's render of that mesh alone."""
    meshes, seg, forest = forest64
    ns, W, H, spp, depth = 32, 512, 512, 8, 4
    mats = np.array([[scenes.MAT_DIFFUSE, 0.7, 0.6, 0.5]], np.float32)
    seg.set_shading(np.zeros(seg.ntri, np.int32), mats)
    cams = []
    for m in meshes[:ns]:
        v = m.reshape(-1, 3)
        c = (v.min(0) + v.max(0)) / 2
        r = float((v.max(0) - v.min(0)).max()) / 2
        cams.append(scenes.Camera(1.0, 40.0, tuple(c + np.array([0.3, 0.5, 3.2]) * r), tuple(c), (0, 1, 0)))
    s = torch.zeros((ns, H, W, 3), dtype=torch.float32, device="cuda")
    q = torch.zeros_like(s)
    torch.cuda.synchronize()
    render_batched(ctx, seg, forest, list(range(ns)), cams, W, H, spp, depth, _lib.MISS_SKY, 3, s.data_ptr(),
                   q.data_ptr())
    ctx.synchronize()
    s, q = s.cpu().numpy(), q.cpu().numpy()
    sizes = [len(m) for m in meshes[:ns]]
    for c in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        print(f"config3 batched scene {c}: {sizes[c]} tris vs oracle", flush=True)
        os_, oq, _ = oracle.render(oracle.Tree(meshes[c]), np.zeros(sizes[c], np.int32), mats, cams[c], W, H, spp,
                                   depth, _lib.MISS_SKY, 3 + c)
        frac = _within_3_sigma(s[c], q[c], os_, oq, spp)
        assert frac < 0.01, (c, frac)
        assert abs(s[c].mean() - os_.mean()) / os_.mean() < 0.005


def test_animation_frame_120_from_graph(ctx, oracle):
    """200k triangles, amplitude 0.02, period 60, seed 5: one frame (animate +
    rebuild) captured once and replayed 120 times."""
    print("config3 animation: 120 graph frames of 200k tris", flush=True)
    pos = scenes.displaced_sphere(200_000, seed=21, radius=1.0)
    rest, mesh = Mesh(ctx, pos), Mesh(ctx, pos)
    bvh = Bvh.lbvh(ctx, mesh)
    frame = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    def one():
        mesh.animate(rest, 0.02, 60.0, 5, frame.data_ptr())
        bvh.rebuild(mesh)

    graph = Graph.capture(ctx, one)
    try:
        for _ in range(120):
            graph.launch()
        ctx.synchronize()
    finally:
        graph.close()
    assert int(frame.cpu()[0]) == 120
    got = mesh.export()
    # the same frame issued eagerly: frame number 119 (the 120th frame)
    m2 = Mesh(ctx, pos)
    f2 = torch.full((1,), 119, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    m2.animate(rest, 0.02, 60.0, 5, f2.data_ptr())
    eager = Bvh.lbvh(ctx, m2)
    assert np.array_equal(m2.export(), got)
    tree = oracle.Tree(got)
    _assert_same_tree(bvh, tree, boxes_exact=False)
    for a, b in zip(bvh.canonical(), eager.canonical()):
        assert np.array_equal(a, b)
    for b in (eager, bvh):
        b.close()
    for m in (m2, mesh, rest):
        m.close()


# ---- differentiation (200k triangles, 512x512, 32 spp, depth 5) -------------------------


@pytest.fixture(scope="module")
def adjoint200k(ctx):
    d = scenes.adjoint_scene(200_000, seed=7, width=512, height=512)
    mesh = Mesh(ctx, d["pos"])
    mesh.set_shading(d["mat_idx"], d["materials"])
    bvh = Bvh.lbvh(ctx, mesh)
    yield d, mesh, bvh
    mesh.bind_params(None, None)
    bvh.close()
    mesh.close()


def test_adjoint_config4_matches_oracle_and_central_differences(ctx, oracle, adjoint200k):
    d, mesh, bvh = adjoint200k
    W, H, spp, depth, seed, miss = 512, 512, 32, 5, 7, _lib.MISS_DEFAULT
    n = len(d["pos"])
    true_alb = torch.from_numpy(d["albedo"]).cuda()
    emi = torch.from_numpy(d["emission"]).cuda()
    alb = torch.full((n, 3), 0.5, dtype=torch.float32, device="cuda")
    s = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    q = torch.empty_like(s)
    torch.cuda.synchronize()

    def mean_image():
        render_device(ctx, mesh, bvh, d["camera"], W, H, spp, depth, miss, seed, s.data_ptr(), q.data_ptr())
        ctx.synchronize()
        return s / spp

    print("config4 adjoint: target and gradient", flush=True)
    mesh.bind_params(true_alb.data_ptr(), emi.data_ptr())
    target = mean_image().double().clone()
    mesh.bind_params(alb.data_ptr(), emi.data_ptr())
    cur = mean_image().double()
    g_img = ((2.0 / cur.numel()) * (cur - target)).float().contiguous()
    ga = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    ge = torch.empty_like(ga)
    torch.cuda.synchronize()
    render_backward(ctx, mesh, bvh, d["camera"], W, H, spp, depth, miss, seed, g_img.data_ptr(), ga.data_ptr(),
                    ge.data_ptr())
    ctx.synchronize()
    ga_h, ge_h = ga.cpu().numpy().astype(np.float64), ge.cpu().numpy().astype(np.float64)

    print("config4 adjoint: FP64 oracle gradient", flush=True)
    _, oga, oge = oracle.render_grad(oracle.Tree(d["pos"]), d["mat_idx"], d["materials"], d["camera"], W, H, spp,
                                     depth, miss, seed, g_img.cpu().numpy().astype(np.float64),
                                     albedo=np.full((n, 3), 0.5, np.float32), emission=d["emission"])
    for got, want in ((ga_h, oga), (ge_h, oge)):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err <= 1e-2, err

    print("config4 adjoint: central differences on 256 triangles", flush=True)

    def loss():
        return float(((mean_image().double() - target) ** 2).mean())

    light = np.nonzero(d["materials"][d["mat_idx"], 0] == scenes.MAT_LIGHT)[0]
    mag = np.abs(ga_h).max(axis=1)
    mag[light] = 0
    # 256 drawn from the 2048 triangles of largest gradient (the gradient is
    # heavy-tailed: few triangles fill much of the image)
    print(f"config4 adjoint: {np.count_nonzero(mag)} triangles with a gradient, "
          f"{np.count_nonzero(mag >= 1e-2 * mag.max())} above 1e-2 and "
          f"{np.count_nonzero(mag >= 1e-3 * mag.max())} above 1e-3 of the largest", flush=True)
    cand = np.argsort(-mag)[:2048]
    tris = np.random.default_rng(17).choice(cand, 256, replace=False)
    h = 1e-3
    worst = 0.0
    picks = [("a", int(t), int(np.argmax(np.abs(ga_h[t])))) for t in tris]
    picks += [("e", int(light[0]), c) for c in range(3)]
    for which, t, c in picks:
        p = alb if which == "a" else emi
        v0 = float(p[t, c])
        vals, steps = [], []
        for sign in (1, -1):
            v = np.float32(v0 + sign * h)
            p[t, c] = float(v)
            steps.append(float(v))
            vals.append(loss())
        p[t, c] = v0
        fd = (vals[0] - vals[1]) / (steps[0] - steps[1])
        want = (ga_h if which == "a" else ge_h)[t, c]
        rel = abs(fd - want) / abs(want)
        worst = max(worst, rel)
        assert rel <= 1e-2, (which, t, c, fd, want)
    print(f"config4 adjoint: worst central-difference error {worst:.2e}", flush=True)
    mesh.bind_params(None, None)


# ---- the large configuration (10.38M triangles) -----------------------------------------


@pytest.fixture(scope="module")
def large(ctx, oracle):
    d = scenes.terrain_scene()
    mesh = Mesh(ctx, d["pos"])
    mesh.set_shading(d["mat_idx"], d["materials"])
    lbvh = Bvh.lbvh(ctx, mesh)
    phr = Bvh.phr(ctx, lbvh)
    otree = oracle.Tree(d["pos"])
    yield d, mesh, lbvh, phr, otree
    mesh.bind_params(None, None)
    phr.close()
    lbvh.close()
    mesh.close()


def test_large_lbvh_matches_oracle(ctx, large):
    d, _, lbvh, _, otree = large
    print(f"config5 LBVH: {len(d['pos'])} tris", flush=True)
    assert len(d["pos"]) == 10_380_418
    _assert_same_tree(lbvh, otree)


def test_large_phr_and_hits_match_oracle(ctx, oracle, large):
    d, _, lbvh, phr, otree = large
    print("config5 PHR vs oracle", flush=True)
    ophr = oracle.PhrTree(otree)
    rec, box = phr.canonical()
    orec, obox, oprims = ophr.canonical()
    assert rec.shape[0] == orec.shape[0]
    assert np.array_equal(rec[:, 0], orec[:, 0]) and np.array_equal(rec[:, 3], orec[:, 2])
    leaf = rec[:, 0] == 1
    assert np.array_equal(rec[leaf, 2] - rec[leaf, 1] + 1, orec[leaf, 1])
    assert np.array_equal(box.astype(np.float64), obox)
    src = phr.phr_source()
    assert np.array_equal(np.concatenate([src[a:b + 1] for a, b in rec[leaf, 1:3]]), oprims)
    assert phr.cost() == pytest.approx(ophr.cost(), rel=1e-12)
    # hits do not depend on the tree (ties aside); the oracle's PHR tree holds
    # leaves of thousands of triangles, so its LBVH answers faster
    print("config5 hits: 200k rays, persistent 128-thread traversal", flush=True)
    pos = d["pos"]
    lo, hi = pos.reshape(-1, 3).min(0), pos.reshape(-1, 3).max(0)
    inc = scenes.incoherent_rays(100_000, seed=91)
    inc[:, :3] = lo + inc[:, :3] * (hi - lo)
    W, H = d["width"], d["height"]
    cam_rays = scenes.camera_rays(d["camera"], W, H).astype(np.float32)[::21][:100_000]
    rays = np.concatenate([inc, cam_rays])
    set_tune(ctx, trav_threads=128, trav_persistent=True)
    try:
        hit, t = phr.closest_hit(rays)
    finally:
        set_tune(ctx)
    ohit, ot = otree.closest_hit(rays)
    assert np.mean(hit == ohit) >= 0.9999
    both = (hit == ohit) & (hit >= 0)
    assert both.sum() > 60_000  # many camera rays see the sky
    assert np.max(np.abs(t[both] - ot[both]) / ot[both]) < 1e-5


def test_large_render_repeatable_and_matches_oracle(ctx, oracle, large):
    d, mesh, _, phr, otree = large
    W, H, spp, depth, seed = d["width"], d["height"], 64, 8, 9
    alb = torch.from_numpy(d["albedo"]).cuda()
    emi = torch.from_numpy(d["emission"]).cuda()
    torch.cuda.synchronize()
    mesh.bind_params(alb.data_ptr(), emi.data_ptr())
    print(f"config5 render {W}x{H}x{spp} twice", flush=True)
    outs = []
    for _ in range(2):
        s = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        q = torch.empty_like(s)
        torch.cuda.synchronize()
        render_device(ctx, mesh, phr, d["camera"], W, H, spp, depth, _lib.MISS_SKY, seed, s.data_ptr(), q.data_ptr())
        ctx.synchronize()
        outs.append((s.cpu().numpy(), q.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][0].mean() > 0
    print("config5 render 96x54x64 vs oracle", flush=True)
    w, h = 96, 54  # the camera's own aspect ratio, 16:9
    cam = d["camera"]
    s, q, _ = render(ctx, mesh, phr, cam, w, h, spp, depth, _lib.MISS_SKY, seed)
    os_, oq, _ = oracle.render(otree, d["mat_idx"], d["materials"], cam, w, h, spp, depth, _lib.MISS_SKY, seed,
                               albedo=d["albedo"], emission=d["emission"])
    assert _within_3_sigma(s, q, os_, oq, spp) < 0.01
    assert abs(s.mean() - os_.mean()) / os_.mean() < 0.01
    mesh.bind_params(None, None)


# ---- the headline tree with the benchmark's traversal variant, and the Sun miss ------------


def test_config2_phr_tree_persistent_128(ctx, oracle):
    print("config2 PHR tree, persistent 128 threads", flush=True)
    pos = scenes.random_triangles(1_000_000, seed=1)
    mesh = Mesh(ctx, pos)
    lbvh = Bvh.lbvh(ctx, mesh)
    phr = Bvh.phr(ctx, lbvh)
    ophr = oracle.PhrTree(oracle.Tree(pos))
    inc = scenes.incoherent_rays(200_000, seed=2)
    coh = scenes.coherent_rays_cfg2(4096, 2048)[::7]
    set_tune(ctx, trav_threads=128, trav_persistent=True)
    try:
        for rays, min_hits in ((inc, 0.8), (coh, 0.3)):
            hit, t = phr.closest_hit(rays)
            ohit, ot = ophr.closest_hit(rays)
            assert np.mean(hit == ohit) >= 0.9999
            both = (hit == ohit) & (hit >= 0)
            assert np.max(np.abs(t[both] - ot[both]) / ot[both]) < 1e-5
            assert np.mean(hit >= 0) >= min_hits
    finally:
        set_tune(ctx)


def test_sun_miss_shader_matches_oracle(ctx, oracle):
    """SunMissShader (render.go:244-246): every escaping path brings (1, .96, .78)."""
    print("render sun miss", flush=True)
    sc = scenes.box_scene(seed=1, slices=40, rings=22, width=160, height=90)
    mesh = Mesh(ctx, sc["pos"])
    mesh.set_shading(sc["mat_idx"], sc["materials"])
    bvh = Bvh.lbvh(ctx, mesh)
    spp = 16
    s, q, last = render(ctx, mesh, bvh, sc["camera"], 160, 90, spp, 5, _lib.MISS_SUN, seed=11)
    os_, oq, olast = oracle.render(oracle.Tree(sc["pos"]), sc["mat_idx"], sc["materials"], sc["camera"], 160, 90,
                                   spp, 5, _lib.MISS_SUN, seed=11)
    assert _within_3_sigma(s, q, os_, oq, spp) < 0.01
    assert abs(s.mean() - os_.mean()) / os_.mean() < 0.01
    # a camera ray through the box's open side that escapes directly sees the sun colour
    sun = np.array([1.0, 0.96, 0.78], np.float32)
    assert np.any(np.all(last == sun, axis=-1))
