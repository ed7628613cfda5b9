"""Batching on the GPU: forests (one LBVH per mesh of a segmented mesh, built
in one pass), batched renders, on-device animation under a captured HIP graph,
and the autotuner.

Bars: every tree of a forest is bit-exact against the FP64 oracle's LBVH of
that mesh alone (keys, permutation, topology, boxes; Cost to 1e-12); scene c of
a batched render equals the single render of its tree with seed + c bit for
bit; graph replays equal the eager calls bit for bit, and a replayed frame's
tree is bit-exact against the oracle built from the exported positions."""
import numpy as np
import pytest

from trace_amd import KEY_M30, KEY_REF63, _lib, scenes
from trace_amd.device import (Bvh, Graph, Mesh, autotune_render, autotune_traversal, get_tune, render,
                              render_batched, render_device, set_tune)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _meshes(sizes, seed):
    out = []
    for i, n in enumerate(sizes):
        p = scenes.random_triangles(n, seed=seed + i, mean_edge=0.01)
        p = p + np.tile(np.array([3.0 * i, -1.0 * i, 0.5 * i]), 3)  # each mesh in its own place
        out.append(p.astype(np.float32).astype(np.float64))
    return out


def _assert_same_tree(view, tree):
    keys, perm = view.sorted()
    okeys, operm = tree.sorted()
    assert np.array_equal(keys, okeys), "Morton keys differ"
    assert np.array_equal(perm, operm), "sort permutation differs"
    rec, box = view.canonical()
    orec, obox = tree.canonical()
    assert rec.shape == orec.shape, (rec.shape, orec.shape)
    assert np.array_equal(rec, orec)
    assert np.array_equal(box.astype(np.float64), obox)
    assert view.cost() == pytest.approx(tree.cost(), rel=1e-12)
    assert view.size() == tree.node_count()


@pytest.mark.parametrize("key_mode", [KEY_REF63, KEY_M30])
def test_forest_trees_match_oracle(ctx, oracle, key_mode):
    sizes = [1, 2, 3, 100, 4095, 4096, 4097, 20_000, 1, 777]
    print(f"forest parity {len(sizes)} meshes mode={key_mode}", flush=True)
    parts = _meshes(sizes, seed=31)
    parts[8] = parts[3][:1].copy()  # a single triangle equal to another mesh's
    dup = np.concatenate([parts[4][:500], parts[4][:500], parts[4][100:200]])
    parts.append(dup)  # equal keys collapse into leaves inside one tree
    mesh = Mesh.segmented(ctx, parts)
    forest = Bvh.lbvh(ctx, mesh, key_mode=key_mode)
    assert forest.info().n_segments == len(parts)
    for i, p in enumerate(parts):
        view = forest.segment(i)
        assert view.n_prims == len(p)
        _assert_same_tree(view, oracle.Tree(p, key_mode=key_mode))
        view.close()


def test_forest_many_segments(ctx, oracle):
    """300 meshes: the segment index needs two radix digits."""
    rng = np.random.default_rng(5)
    sizes = rng.integers(1, 60, 300).tolist()
    print(f"forest parity 300 meshes, {sum(sizes)} tris", flush=True)
    parts = _meshes(sizes, seed=1000)
    forest = Bvh.lbvh(ctx, Mesh.segmented(ctx, parts))
    for i in (0, 1, 57, 255, 256, 257, 299):
        v = forest.segment(i)
        _assert_same_tree(v, oracle.Tree(parts[i]))
        v.close()


def test_forest_closest_hit_per_tree(ctx, oracle):
    print("forest closest hit", flush=True)
    parts = _meshes([5000, 20_000, 3], seed=41)
    forest = Bvh.lbvh(ctx, Mesh.segmented(ctx, parts))
    for i, p in enumerate(parts):
        lo = p.reshape(-1, 3).min(0)
        hi = p.reshape(-1, 3).max(0)
        rays = scenes.incoherent_rays(20_000, seed=50 + i)
        rays[:, :3] = lo + rays[:, :3] * (hi - lo)  # origins inside this mesh's box
        v = forest.segment(i)
        hit, t = v.closest_hit(rays)
        ohit, ot = oracle.Tree(p).closest_hit(rays)
        assert np.mean(hit == ohit) >= 0.9999
        both = (hit == ohit) & (hit >= 0)
        rel = np.abs(t[both] - ot[both]) / np.maximum(np.abs(ot[both]), 1e-30)
        assert rel.size == 0 or rel.max() <= 1e-5
        assert hit.max() < len(p)  # indices are local to the segment
        v.close()


def test_forest_rejects_whole_forest_queries(ctx):
    forest = Bvh.lbvh(ctx, Mesh.segmented(ctx, _meshes([10, 20], seed=3)))
    with pytest.raises(_lib.TraceError):
        forest.closest_hit(scenes.incoherent_rays(10))
    with pytest.raises(_lib.TraceError):
        forest.segment(2)
    with pytest.raises(_lib.TraceError):
        Bvh.phr(ctx, forest)


def _shaded_batch(ctx, n_meshes=5):
    parts = [scenes.displaced_sphere(n, seed=60 + i, center=(2.5 * i, 0.0, 0.0), radius=0.8)
             for i, n in enumerate([3000, 800, 12_000, 200, 5000][:n_meshes])]
    mesh = Mesh.segmented(ctx, parts)
    mats = np.array([[scenes.MAT_DIFFUSE, 0.7, 0.5, 0.3], [scenes.MAT_DIFFUSE, 0.2, 0.8, 0.4]], np.float32)
    mi = np.concatenate([np.full(len(p), i % 2, np.int32) for i, p in enumerate(parts)])
    mi[::7] = 1 - mi[::7]
    mesh.set_shading(mi, mats)
    cams = [scenes.Camera(1.0, 50.0, (2.5 * i + 0.3, 0.6, 2.8), (2.5 * i, 0.0, 0.0), (0, 1, 0))
            for i in range(len(parts))]
    return parts, mesh, cams


@pytest.mark.parametrize("render_batch", [1 << 22, 1])
def test_batched_render_equals_single_renders(ctx, render_batch):
    print(f"batched render render_batch={render_batch}", flush=True)
    parts, mesh, cams = _shaded_batch(ctx)
    forest = Bvh.lbvh(ctx, mesh)
    seg_ids = [3, 0, 2, 4, 1, 3]
    W, H, spp, depth, seed = 48, 40, 3, 4, 17
    set_tune(ctx, render_batch=render_batch)
    try:
        ns = len(seg_ids)
        s = torch.zeros((ns, H, W, 3), dtype=torch.float32, device="cuda")
        q = torch.zeros_like(s)
        last = torch.zeros_like(s)
        torch.cuda.synchronize()
        render_batched(ctx, mesh, forest, seg_ids, [cams[i] for i in seg_ids], W, H, spp, depth,
                       _lib.MISS_SKY, seed, s.data_ptr(), q.data_ptr(), last.data_ptr())
        ctx.synchronize()
        s, q, last = s.cpu().numpy(), q.cpu().numpy(), last.cpu().numpy()
        for c, seg in enumerate(seg_ids):
            v = forest.segment(seg)
            s1, q1, l1 = render(ctx, mesh, v, cams[seg], W, H, spp, depth, _lib.MISS_SKY, seed + c)
            assert np.array_equal(s[c], s1), f"scene {c}: sums differ"
            assert np.array_equal(q[c], q1), f"scene {c}: sums of squares differ"
            assert np.array_equal(last[c], l1), f"scene {c}: last sample differs"
            assert s1.max() > 0
            v.close()
    finally:
        set_tune(ctx)


def test_batched_render_argument_checks(ctx):
    parts, mesh, cams = _shaded_batch(ctx, 2)
    forest = Bvh.lbvh(ctx, mesh)
    s = torch.zeros((1, 8, 8, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.TraceError):
        render_batched(ctx, mesh, forest, [2], [cams[0]], 8, 8, 1, 1, 0, 1, s.data_ptr(), s.data_ptr())
    with pytest.raises(_lib.TraceError):  # a forest renders through its views
        render(ctx, mesh, forest, cams[0], 8, 8, 1, 1, 0, 1)


def _outward_f32(box64):
    lo, hi = box64[:, :3], box64[:, 3:]
    flo, fhi = lo.astype(np.float32), hi.astype(np.float32)
    flo = np.where(flo.astype(np.float64) > lo, np.nextafter(flo, np.float32(-np.inf)), flo)
    fhi = np.where(fhi.astype(np.float64) < hi, np.nextafter(fhi, np.float32(np.inf)), fhi)
    return np.concatenate([flo, fhi], axis=1)


def _anim_scene(ctx, ntri):
    pos = scenes.displaced_sphere(ntri, seed=71, radius=0.5)
    mats = np.array([[scenes.MAT_DIFFUSE, 0.6, 0.6, 0.6]], np.float32)
    rest = Mesh(ctx, pos)
    mesh = Mesh(ctx, pos)
    mesh.set_shading(np.zeros(len(pos), np.int32), mats)
    return pos, rest, mesh


def test_animation_graph_matches_eager(ctx, oracle):
    """Frames recorded once in a HIP graph (animate, rebuild, closest hit, render)
    and replayed equal the same calls issued eagerly, frame by frame."""
    print("animation graph vs eager", flush=True)
    ntri, frames, amp, period, aseed = 20_000, 4, 0.02, 30.0, 9
    W, H = 32, 24
    cam = scenes.Camera(W / H, 45.0, (0.2, 0.3, 1.8), (0, 0, 0), (0, 1, 0))
    rays = torch.from_numpy(scenes.incoherent_rays(30_000, seed=72)).cuda()

    def setup():
        pos, rest, mesh = _anim_scene(ctx, ntri)
        bvh = Bvh.lbvh(ctx, mesh)
        st = {
            "rest": rest, "mesh": mesh, "bvh": bvh,
            "frame": torch.zeros(1, dtype=torch.int32, device="cuda"),
            "hit": torch.empty(len(rays), dtype=torch.int32, device="cuda"),
            "t": torch.empty(len(rays), dtype=torch.float32, device="cuda"),
            "sum": torch.empty((H, W, 3), dtype=torch.float32, device="cuda"),
            "sq": torch.empty((H, W, 3), dtype=torch.float32, device="cuda"),
        }
        torch.cuda.synchronize()
        return st

    def step(st):
        st["mesh"].animate(st["rest"], amp, period, aseed, st["frame"].data_ptr())
        st["bvh"].rebuild(st["mesh"])
        st["bvh"].closest_hit_device(rays.data_ptr(), len(rays), st["hit"].data_ptr(), st["t"].data_ptr())
        render_device(ctx, st["mesh"], st["bvh"], cam, W, H, 2, 3, _lib.MISS_SKY, 5, st["sum"].data_ptr(),
                      st["sq"].data_ptr())

    def snap(st):
        ctx.synchronize()
        return [st[k].cpu().numpy().copy() for k in ("hit", "t", "sum", "sq")]

    eager = setup()
    want = []
    for _ in range(frames):
        step(eager)
        want.append(snap(eager))

    g = setup()
    # size the scratch arena for every call first: it may not move under a graph
    render_device(ctx, g["mesh"], g["bvh"], cam, W, H, 2, 3, _lib.MISS_SKY, 5, g["sum"].data_ptr(),
                  g["sq"].data_ptr())
    ctx.synchronize()
    graph = Graph.capture(ctx, lambda: step(g))
    try:
        for f in range(frames):
            graph.launch()
            got = snap(g)
            for name, a, b in zip(("hit", "t", "sum", "sq"), got, want[f]):
                assert np.array_equal(a, b), f"frame {f}: {name} differs between graph and eager"
        assert int(g["frame"].cpu()[0]) == frames
        # the replayed frame's tree is the LBVH of the exported positions
        pos = g["mesh"].export()
        keys, perm = g["bvh"].sorted()
        tree = oracle.Tree(pos)
        okeys, operm = tree.sorted()
        assert np.array_equal(keys, okeys) and np.array_equal(perm, operm)
        rec, box = g["bvh"].canonical()
        orec, obox = tree.canonical()
        assert np.array_equal(rec, orec)
        # animated vertices are not FP32-exact: the FP32 boxes are the FP64 boxes
        # rounded outwards
        assert np.array_equal(box, _outward_f32(obox))
        # the scratch arena cannot grow while the graph is alive: a render whose
        # 2^29 paths of state no earlier call of the session has needed
        old = get_tune(ctx)
        set_tune(ctx, old.trav_threads, old.render_threads, 1 << 29, bool(old.trav_persistent),
                 bool(old.render_persistent), old.sort_bits, old.sort_keys_per_thread)
        big = torch.empty((2048, 2048, 3), dtype=torch.float32, device="cuda")
        try:
            with pytest.raises(_lib.TraceError, match="scratch arena"):
                render_device(ctx, g["mesh"], g["bvh"], cam, 2048, 2048, 128, 3, _lib.MISS_SKY, 5, big.data_ptr(),
                              big.data_ptr())
        finally:
            set_tune(ctx, old.trav_threads, old.render_threads, old.render_batch, bool(old.trav_persistent),
                     bool(old.render_persistent), old.sort_bits, old.sort_keys_per_thread)
    finally:
        graph.close()


def test_animation_displacement_is_rigid_per_vertex(ctx):
    """Every vertex moves along one direction by at most the amplitude, and
    vertices shared between triangles stay shared."""
    print("animation displacement", flush=True)
    pos, rest, mesh = _anim_scene(ctx, 5000)
    frame = torch.tensor([7], dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    mesh.animate(rest, 0.02, 30.0, 9, frame.data_ptr())
    out = mesh.export()
    assert int(frame.cpu()[0]) == 8
    d = (out - pos).reshape(-1, 3)
    mag = np.linalg.norm(d, axis=1)
    assert mag.max() <= 0.02 * (1 + 1e-12) and mag.max() > 0.01
    big = d[mag > 1e-3]
    u = big / np.linalg.norm(big, axis=1, keepdims=True)
    assert np.all(np.abs(np.abs(u @ u[0]) - 1.0) < 1e-9)
    v_rest = pos.reshape(-1, 3)
    v_out = out.reshape(-1, 3)
    _, first, inv = np.unique(v_rest, axis=0, return_index=True, return_inverse=True)
    assert np.array_equal(v_out, v_out[first][inv.ravel()])


def test_graph_api_errors(ctx):
    with pytest.raises(_lib.TraceError):
        ctx.check(ctx.lib.trace_graph_end(ctx.h, _lib.C.byref(_lib.C.c_void_p())))


def test_autotune_keeps_results(ctx):
    print("autotune", flush=True)
    pos = scenes.random_triangles(50_000, seed=81)
    mesh = Mesh(ctx, pos)
    bvh = Bvh.lbvh(ctx, mesh)
    rays_h = scenes.incoherent_rays(100_000, seed=82)
    rays = torch.from_numpy(rays_h).cuda()
    torch.cuda.synchronize()
    ms = autotune_traversal(ctx, bvh, rays.data_ptr(), len(rays))
    assert ms.shape == (9,) and np.all(ms > 0)
    t = get_tune(ctx)
    assert t.tuned == 1 and t.trav_threads in (64, 128, 256) and t.trav_persistent in (0, 1)
    assert t.trav_order in (0, 1) and (t.trav_persistent or not t.trav_order)
    ref = None
    for persistent, order in ((False, False), (True, False), (True, True)):
        for threads in (64, 128, 256):
            set_tune(ctx, trav_threads=threads, trav_persistent=persistent, trav_order=order)
            hit, tt = bvh.closest_hit(rays_h)
            if ref is None:
                ref = (hit, tt)
            assert np.array_equal(hit, ref[0]) and np.array_equal(tt, ref[1])

    sc = scenes.box_scene(seed=3, slices=30, rings=20, width=64, height=36)
    smesh = Mesh(ctx, sc["pos"])
    smesh.set_shading(sc["mat_idx"], sc["materials"])
    sbvh = Bvh.lbvh(ctx, smesh)
    set_tune(ctx)
    rms = autotune_render(ctx, smesh, sbvh, sc["camera"], 64, 36, 2, 3, _lib.MISS_DEFAULT, 3)
    assert rms.shape == (9,) and np.all(rms > 0)
    t = get_tune(ctx)
    assert t.render_threads in (64, 128, 256) and t.render_batch in (1 << 22, 1 << 24, 1 << 26, 1 << 27)
    imgs = []
    for persistent in (False, True):
        for threads in (64, 128, 256):
            set_tune(ctx, render_threads=threads, render_persistent=persistent)
            imgs.append(render(ctx, smesh, sbvh, sc["camera"], 64, 36, 2, 3, _lib.MISS_DEFAULT, 3))
    for other in imgs[1:]:
        for a, b in zip(imgs[0], other):
            assert np.array_equal(a, b)
    set_tune(ctx)
    with pytest.raises(_lib.TraceError):
        set_tune(ctx, trav_threads=96)
