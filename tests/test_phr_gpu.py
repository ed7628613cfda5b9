"""GPU parity of PHR refinement against the FP64 oracle.

Both sides take the initial cut in FIFO order and follow phr.go step by step,
so the tree is compared exactly: node kinds and arities in preorder, leaf
triangle lists (as LBVH sorted positions), boxes, and Cost()."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Graph, Mesh

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _compare(ctx, oracle, pos, alpha=0.5, delta=6.0, max_cut=2500, branching=2):
    mesh = Mesh(ctx, pos)
    aux = Bvh.lbvh(ctx, mesh)
    phr = Bvh.phr(ctx, aux, alpha, delta, branching, max_cut)
    otree = oracle.Tree(pos)
    ophr = oracle.PhrTree(otree, alpha, delta, branching, max_cut)
    rec, box = phr.canonical()
    orec, obox, oprims = ophr.canonical()
    assert rec.shape[0] == orec.shape[0], (rec.shape, orec.shape)
    assert np.array_equal(rec[:, 0], orec[:, 0]), "node kinds differ"
    assert np.array_equal(rec[:, 3], orec[:, 2]), "arities differ"
    leaf = rec[:, 0] == 1
    assert np.array_equal(rec[leaf, 2] - rec[leaf, 1] + 1, orec[leaf, 1]), "leaf sizes differ"
    assert np.array_equal(box.astype(np.float64), obox), "boxes differ"
    # leaf triangles, in leaf order, as LBVH sorted positions
    src = phr.phr_source()
    gprims = np.concatenate([src[a:b + 1] for a, b in rec[leaf, 1:3]])
    assert np.array_equal(gprims, oprims), "leaf triangle lists differ"
    assert phr.cost() == pytest.approx(ophr.cost(), rel=1e-12)
    assert phr.size() == rec.shape[0]
    # every triangle is in exactly one leaf and perm is a bijection
    perm = phr.perm()
    assert np.array_equal(np.sort(perm), np.arange(len(pos)))
    # the PHR order maps back to the LBVH order through the source positions
    _, lperm = aux.sorted()
    assert np.array_equal(perm, lperm[src])
    return mesh, aux, phr, otree, ophr


@pytest.mark.parametrize("n,seed", [(2, 1), (7, 2), (100, 3), (5000, 4), (50_000, 5)])
def test_phr_matches_oracle(ctx, oracle, n, seed):
    print(f"phr parity n={n}", flush=True)
    _compare(ctx, oracle, scenes.random_triangles(n, seed=seed))


def test_phr_hq_settings_and_cap(ctx, oracle):
    """PHR_HQ (alpha 0.55, delta 9, benchmark_test.go:16-17): the 2500 cap binds
    and cuts above 4096 members run through the global-memory sort path."""
    print("phr hq", flush=True)
    _, _, phr, _, ophr = _compare(ctx, oracle, scenes.random_triangles(200_000, seed=6), 0.55, 9.0)
    assert phr.info().phr_initial_cut == ophr.initial_cut() == 2500


def test_phr_small_cap(ctx, oracle):
    print("phr small cap", flush=True)
    _compare(ctx, oracle, scenes.random_triangles(20_000, seed=7), max_cut=37)


def test_phr_box_scene(ctx, oracle):
    print("phr box scene", flush=True)
    sc = scenes.box_scene(seed=1)
    _compare(ctx, oracle, sc["pos"])


def test_phr_config2_full(ctx, oracle):
    """1M triangles, default builder (alpha 0.5, delta 6, branching 2)."""
    print("phr 1M", flush=True)
    _, _, phr, otree, ophr = _compare(ctx, oracle, scenes.random_triangles(1_000_000, seed=1))
    assert phr.cost() < otree.cost()  # PHR improves the LBVH's SAH cost


def test_phr_traversal_and_render(ctx, oracle):
    """Closest hit on the PHR tree equals the oracle's LBVH answer (the hit does
    not depend on the tree), and a render over it matches the oracle."""
    print("phr traversal", flush=True)
    pos = scenes.random_triangles(100_000, seed=8, mean_edge=0.01)
    mesh = Mesh(ctx, pos)
    aux = Bvh.lbvh(ctx, mesh)
    phr = Bvh.phr(ctx, aux)
    rays = scenes.incoherent_rays(100_000, seed=9)
    hit, t = phr.closest_hit(rays)
    ohit, ot = oracle.Tree(pos).closest_hit(rays)
    assert np.mean(hit == ohit) >= 0.9999
    both = (hit == ohit) & (hit >= 0)
    assert np.max(np.abs(t[both] - ot[both]) / ot[both]) < 1e-5
    sc = scenes.box_scene(seed=1, slices=40, rings=22, width=96, height=54)
    m2 = Mesh(ctx, sc["pos"])
    m2.set_shading(sc["mat_idx"], sc["materials"])
    p2 = Bvh.phr(ctx, Bvh.lbvh(ctx, m2))
    from trace_amd.device import render

    s, q, _ = render(ctx, m2, p2, sc["camera"], 96, 54, 8, 5, _lib.MISS_SKY, seed=4)
    os_, oq, _ = oracle.render(oracle.Tree(sc["pos"]), sc["mat_idx"], sc["materials"], sc["camera"], 96, 54, 8, 5,
                               _lib.MISS_SKY, seed=4)
    from test_render_gpu import within_3_sigma

    frac, _ = within_3_sigma(s, q, os_, oq, 8)
    assert frac < 0.01


def test_phr_terrain_big_leaves_split(ctx, oracle):
    """On a heightfield, PHR's makeLeaf flattens whole LBVH subtrees into
    leaves of hundreds of triangles (phr.go:313-352); the tree still equals
    the oracle's.  The 4-wide traversal tree splits those leaves: closest
    hits equal the oracle's, and triangle tests per ray stay close to the
    LBVH's instead of growing with the leaf size."""
    print("phr terrain big leaves", flush=True)
    d = scenes.terrain_scene(n=384, n_spheres=200, width=192, height=108)
    pos = d["pos"]
    _, aux, phr, otree, _ = _compare(ctx, oracle, pos)
    rec, _ = phr.canonical()
    leaf = rec[:, 0] == 1
    assert (rec[leaf, 2] - rec[leaf, 1] + 1).max() > 64  # the premise: big PHR leaves
    lo, hi = pos.reshape(-1, 3).min(0), pos.reshape(-1, 3).max(0)
    inc = scenes.incoherent_rays(20_000, seed=31).astype(np.float32)
    inc[:, :3] = lo + inc[:, :3] * (hi - lo)
    rays = np.concatenate([scenes.camera_rays(d["camera"], 192, 108).astype(np.float32), inc])
    hit, t = phr.closest_hit(rays)
    ohit, ot = otree.closest_hit(rays)
    assert np.mean(hit == ohit) >= 0.9999
    both = (hit == ohit) & (hit >= 0)
    assert both.sum() > 10_000
    assert np.max(np.abs(t[both] - ot[both]) / ot[both]) < 1e-5
    _, tp = phr.counts(rays)
    _, tl = aux.counts(rays)
    assert tp.mean() <= 2.0 * tl.mean() + 4.0, (tp.mean(), tl.mean())


@pytest.mark.parametrize("alpha,delta", [(0.5, 6.0), (0.4, 9.0), (0.55, 7.0)])
def test_build_with_cost_matches_oracle(ctx, oracle, alpha, delta):
    """PhrBuilder.BuildWithCost (phr.go:154-255): the build cost (initial cut +
    every cut handed to SweepSAH) equals the oracle's, and so does Cost()."""
    print(f"phr build cost alpha={alpha} delta={delta}", flush=True)
    pos = scenes.random_triangles(50_000, seed=23)
    aux = Bvh.lbvh(ctx, Mesh(ctx, pos))
    phr, cost = Bvh.phr_with_cost(ctx, aux, alpha, delta)
    ref = oracle.PhrTree(oracle.Tree(pos), alpha, delta)
    assert cost == ref.build_cost()
    assert phr.info().phr_build_cost == cost
    assert phr.cost() == pytest.approx(ref.cost(), rel=1e-12)


def test_build_with_cost_small_trees(ctx, oracle):
    for n in (1, 2, 3):
        pos = scenes.random_triangles(n, seed=24)
        _, cost = Bvh.phr_with_cost(ctx, Bvh.lbvh(ctx, Mesh(ctx, pos)))
        assert cost == oracle.PhrTree(oracle.Tree(pos)).build_cost(), n


def test_grid_optimizer_matches_oracle(ctx, oracle):
    """gridOptimizer.OptimizedPHRparams over the default 4 x 4 grid at 1920x1080:
    every evaluation and the chosen (alpha, delta) equal the oracle's."""
    from oracle.pyoracle import grid_optimize
    from trace_amd.device import phr_optimize_grid

    print("phr grid optimizer", flush=True)
    pos = scenes.random_triangles(20_000, seed=25)
    aux = Bvh.lbvh(ctx, Mesh(ctx, pos))
    a, d, sah, build = phr_optimize_grid(ctx, aux)
    oa, od, ev = grid_optimize(oracle.Tree(pos), (0.4, 0.45, 0.5, 0.55), (6.0, 7.0, 8.0, 9.0), 1920, 1080)
    assert np.array_equal(build, np.array([e[1] for e in ev]))
    assert np.allclose(sah, np.array([e[0] for e in ev]), rtol=1e-12, atol=0)
    assert (a, d) == (oa, od)
    # a frame large enough that the SAH term decides
    a2, d2, _, _ = phr_optimize_grid(ctx, aux, width=100_000, height=100_000)
    oa2, od2, _ = grid_optimize(oracle.Tree(pos), (0.4, 0.45, 0.5, 0.55), (6.0, 7.0, 8.0, 9.0), 100_000, 100_000)
    assert (a2, d2) == (oa2, od2)


def test_bayes_optimizer_matches_oracle(ctx, oracle):
    """bayesianOptimizer.OptimizedPHRparams: the normaliser builds and, at every
    point the search evaluated, the build cost, Cost() and objective equal the
    oracle's; the result is the evaluated minimum and the same seed repeats the
    run.  Which points the reference's third-party search (changkun.de/x/bo)
    would try is parity unpinned."""
    from oracle.pyoracle import bayes_objective
    from trace_amd.device import phr_optimize_bayes

    print("phr bayes optimizer", flush=True)
    pos = scenes.random_triangles(20_000, seed=26)
    aux = Bvh.lbvh(ctx, Mesh(ctx, pos))
    ar, dr = (0.4, 0.55), (6.0, 9.0)
    a, d, tr = phr_optimize_bayes(ctx, aux, ar, dr, width=1920, height=1080, rounds=8, random_rounds=3, seed=4)
    tree = oracle.Tree(pos)
    lo = oracle.PhrTree(tree, ar[0], dr[0])
    hi = oracle.PhrTree(tree, ar[1], dr[1])
    assert tuple(tr[0, :2]) == (ar[0], dr[0]) and tuple(tr[1, :2]) == (ar[1], dr[1])
    assert tr[0, 3] == lo.build_cost() and tr[1, 3] == hi.build_cost()
    assert np.allclose(tr[:2, 2], [lo.cost(), hi.cost()], rtol=1e-12, atol=0)
    assert np.all(np.isnan(tr[:2, 4]))
    ev = tr[2:]
    assert np.all((ev[:, 0] >= ar[0]) & (ev[:, 0] <= ar[1]) & (ev[:, 1] >= dr[0]) & (ev[:, 1] <= dr[1]))
    assert len({(r[0], r[1]) for r in ev}) > 4  # the search moves
    for al, de, sah, build, y in ev:
        oy, osah, obuild = bayes_objective(tree, al, de, ar, dr, 1920, 1080)
        assert build == obuild, (al, de)
        assert sah == pytest.approx(osah, rel=1e-12, abs=0)
        assert y == pytest.approx(oy, rel=1e-12, abs=0)
    k = int(np.argmin(ev[:, 4]))
    assert (a, d) == (ev[k, 0], ev[k, 1])
    a2, d2, tr2 = phr_optimize_bayes(ctx, aux, ar, dr, width=1920, height=1080, rounds=8, random_rounds=3, seed=4)
    assert (a2, d2) == (a, d) and np.array_equal(tr2[:, :4], tr[:, :4])


def test_phr_rejects_unsupported(ctx):
    pos = scenes.random_triangles(100, seed=1)
    aux = Bvh.lbvh(ctx, Mesh(ctx, pos))
    for bf in (1, 9):  # this build takes branching factors 2 to 8
        with pytest.raises(_lib.TraceError):
            Bvh.phr(ctx, aux, branching=bf)
    with pytest.raises(_lib.TraceError):
        Bvh.phr(ctx, aux, max_cut=0)


def _same_tree(a, b):
    ra, ba = a.canonical()
    rb, bb = b.canonical()
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)
    assert np.array_equal(a.phr_source(), b.phr_source())
    assert np.array_equal(a.perm(), b.perm())
    assert a.cost() == b.cost() and a.size() == b.size()


def test_phr_rebuild_in_place(ctx, oracle):
    """trace_phr_rebuild gives the tree a fresh trace_phr_build gives, whatever the
    tree held before: other settings, other branching factors (binary after n-ary
    and back), a one-node initial cut (the LBVH itself) and another mesh."""
    print("phr rebuild in place", flush=True)
    pos = scenes.random_triangles(30_000, seed=31)
    aux = Bvh.lbvh(ctx, Mesh(ctx, pos))
    tree = Bvh.phr(ctx, aux)
    cases = [(0.55, 9.0, 2, 2500), (0.5, 6.0, 4, 2500), (0.4, 7.0, 2, 100), (0.5, 6.0, 8, 2500),
             (0.5, 6.0, 2, 1), (0.45, 8.0, 3, 2500), (0.5, 6.0, 2, 2500)]
    for alpha, delta, bf, cap in cases:
        tree.phr_rebuild(aux, alpha, delta, bf, cap)
        fresh = Bvh.phr(ctx, aux, alpha, delta, bf, cap)
        _same_tree(tree, fresh)
        assert tree.info().arity == fresh.info().arity
        fresh.close()
    # the last case against the oracle, and the traversal of the rebuilt tree
    rec, box = tree.canonical()
    orec, obox, _ = oracle.PhrTree(oracle.Tree(pos), 0.5, 6.0, 2, 2500).canonical()
    assert np.array_equal(rec[:, 0], orec[:, 0]) and np.array_equal(box.astype(np.float64), obox)
    rays = scenes.incoherent_rays(20_000, seed=32)
    hit, t = tree.closest_hit(rays)
    ohit, ot = oracle.Tree(pos).closest_hit(rays)
    assert np.mean(hit == ohit) >= 0.9999
    # another, larger mesh into the same storage
    pos2 = scenes.random_triangles(60_000, seed=33)
    aux2 = Bvh.lbvh(ctx, Mesh(ctx, pos2))
    tree.phr_rebuild(aux2)
    fresh = Bvh.phr(ctx, aux2)
    _same_tree(tree, fresh)


def test_phr_rebuild_refusals(ctx):
    pos = scenes.random_triangles(1000, seed=34)
    aux = Bvh.lbvh(ctx, Mesh(ctx, pos))
    tree = Bvh.phr(ctx, aux)
    with pytest.raises(_lib.TraceError):  # the target is an LBVH
        aux.phr_rebuild(aux)
    with pytest.raises(_lib.TraceError):  # the auxiliary tree is a PHR tree
        tree.phr_rebuild(tree)
    for bf in (1, 9):
        with pytest.raises(_lib.TraceError):
            tree.phr_rebuild(aux, branching=bf)
    # a refused call leaves the tree as it was
    fresh = Bvh.phr(ctx, aux)
    _same_tree(tree, fresh)


def _outward_f32(box64):
    """FP64 boxes rounded outwards to FP32 (the device's boxes of animated,
    not FP32-exact, vertices)."""
    lo, hi = box64[:, :3], box64[:, 3:]
    flo, fhi = lo.astype(np.float32), hi.astype(np.float32)
    flo = np.where(flo.astype(np.float64) > lo, np.nextafter(flo, np.float32(-np.inf)), flo)
    fhi = np.where(fhi.astype(np.float64) < hi, np.nextafter(fhi, np.float32(np.inf)), fhi)
    return np.concatenate([flo, fhi], axis=1)


def test_phr_rebuild_in_a_graph(ctx, oracle):
    """A frame recorded once in a HIP graph -- animate the mesh, rebuild its LBVH,
    refine it into the PHR tree's storage, trace rays -- replays to the trees and
    hits the same calls give eagerly, frame by frame; the last replayed frame's
    PHR tree is the oracle's for the exported positions, and the header's levels
    and output nodes follow the replays."""
    print("phr rebuild in a graph", flush=True)
    ntri, frames, amp, period, aseed = 20_000, 4, 0.02, 30.0, 9
    pos = scenes.displaced_sphere(ntri, seed=71, radius=0.5)
    rays = torch.from_numpy(scenes.incoherent_rays(30_000, seed=74)).cuda()

    def setup():
        rest, mesh = Mesh(ctx, pos), Mesh(ctx, pos)
        aux = Bvh.lbvh(ctx, mesh)
        tree = Bvh.phr(ctx, aux)
        tree.phr_rebuild(aux)  # the eager rebuild a captured one repeats
        st = {"rest": rest, "mesh": mesh, "aux": aux, "tree": tree,
              "frame": torch.zeros(1, dtype=torch.int32, device="cuda"),
              "hit": torch.empty(len(rays), dtype=torch.int32, device="cuda"),
              "t": torch.empty(len(rays), dtype=torch.float32, device="cuda")}
        torch.cuda.synchronize()
        return st

    def step(st):
        st["mesh"].animate(st["rest"], amp, period, aseed, st["frame"].data_ptr())
        st["aux"].rebuild(st["mesh"])
        st["tree"].phr_rebuild(st["aux"])
        st["tree"].closest_hit_device(rays.data_ptr(), len(rays), st["hit"].data_ptr(), st["t"].data_ptr())

    def snap(st):
        ctx.synchronize()
        rec, box = st["tree"].canonical()
        return [st["hit"].cpu().numpy().copy(), st["t"].cpu().numpy().copy(), rec, box, st["tree"].perm(),
                st["tree"].phr_source(), st["tree"].cost(), st["tree"].info().phr_levels]

    eager = setup()
    want = []
    for _ in range(frames):
        step(eager)
        want.append(snap(eager))

    g = setup()
    # other settings than the eager rebuild's, and branching 4, are refused while capturing
    for kw in ({"alpha": 0.55}, {"branching": 4}):
        with pytest.raises(_lib.TraceError):
            Graph.capture(ctx, lambda: g["tree"].phr_rebuild(g["aux"], **kw))
    graph = Graph.capture(ctx, lambda: step(g))
    try:
        for f in range(frames):
            graph.launch()
            got = snap(g)
            for k, (a, b) in enumerate(zip(got, want[f])):
                assert np.array_equal(np.asarray(a), np.asarray(b)), f"frame {f}: item {k} differs"
        assert int(g["frame"].cpu()[0]) == frames
        # the replayed frame's tree against the oracle's PHR of the exported positions
        opos = g["mesh"].export()
        rec, box = g["tree"].canonical()
        orec, obox, oprims = oracle.PhrTree(oracle.Tree(opos), 0.5, 6.0, 2, 2500).canonical()
        assert rec.shape[0] == orec.shape[0]
        assert np.array_equal(rec[:, 0], orec[:, 0]) and np.array_equal(rec[:, 3], orec[:, 2])
        leaf = rec[:, 0] == 1
        assert np.array_equal(rec[leaf, 2] - rec[leaf, 1] + 1, orec[leaf, 1])
        assert np.array_equal(box, _outward_f32(obox))
        src = g["tree"].phr_source()
        assert np.array_equal(np.concatenate([src[a:b + 1] for a, b in rec[leaf, 1:3]]), oprims)
        ohit, _ = oracle.Tree(opos).closest_hit(rays.cpu().numpy())
        assert np.mean(g["hit"].cpu().numpy() == ohit) >= 0.9999
    finally:
        graph.close()


def test_phr_rebuild_graph_out_of_bounds(ctx):
    """A replay whose frame leaves the recorded bounds -- here every triangle the
    same, so the LBVH is one leaf and the initial cut one node, which an eager
    rebuild answers by attaching the LBVH as is -- reports TRACE_ERR_BOUND at the
    next synchronisation and leaves the previous frame's tree in place; the next
    frame inside the bounds replays normally."""
    print("phr rebuild graph out of bounds", flush=True)
    n = 5000
    pos = scenes.random_triangles(n, seed=75)
    same = np.repeat(pos[:1], n, axis=0)
    mesh = Mesh(ctx, pos)
    aux = Bvh.lbvh(ctx, mesh)
    tree = Bvh.phr(ctx, aux)
    tree.phr_rebuild(aux)
    fresh = Bvh.phr(ctx, aux)
    ctx.synchronize()
    graph = Graph.capture(ctx, lambda: (aux.rebuild(mesh), tree.phr_rebuild(aux)))
    try:
        graph.launch()
        ctx.synchronize()
        _same_tree(tree, fresh)
        mesh.update(same)
        graph.launch()
        with pytest.raises(_lib.TraceError, match="bounds"):
            ctx.synchronize()
        _same_tree(tree, fresh)  # the previous frame's tree
        mesh.update(pos)
        graph.launch()
        ctx.synchronize()
        _same_tree(tree, fresh)
    finally:
        graph.close()
    torch.cuda.synchronize()


# ---- branching factors above 2 (phr.go:92-152, scene.go:32-36) ---------------------------


@pytest.mark.parametrize("bf", [3, 4, 8])
@pytest.mark.parametrize("n,seed", [(2, 1), (7, 2), (100, 3), (5000, 4), (50_000, 5)])
def test_phr_nary_matches_oracle(ctx, oracle, n, seed, bf):
    """n-ary trees: node kinds and child counts in preorder, leaf lists, FP64
    boxes and Cost() equal the oracle's; BuildWithCost's counter too."""
    print(f"phr n-ary parity n={n} bf={bf}", flush=True)
    _, aux, phr, _, ophr = _compare(ctx, oracle, scenes.random_triangles(n, seed=seed), branching=bf)
    rec, _ = phr.canonical()
    assert rec[:, 3].max() <= bf
    if n >= 5000:
        assert rec[:, 3].max() == bf
    info = phr.info()
    assert info.arity == bf
    cost = Bvh.phr_with_cost(ctx, aux, 0.5, 6.0, bf)[1]
    assert cost == ophr.build_cost()


@pytest.mark.parametrize("bf", [3, 4])
def test_phr_nary_settings(ctx, oracle, bf):
    """The HQ setting (the 2500 cap binds), a small cap, no refinement beyond
    the initial cut and full refinement, each with branching 3 and 4."""
    print(f"phr n-ary settings bf={bf}", flush=True)
    pos = scenes.random_triangles(60_000, seed=6)
    for alpha, delta, cap in ((0.55, 9.0, 2500), (0.5, 6.0, 37), (0.0, 2.0, 2500), (0.0, 30.0, 2500),
                              (1.0, 0.5, 2500)):
        _compare(ctx, oracle, pos, alpha, delta, cap, branching=bf)


def test_phr_nary_config2_full(ctx, oracle):
    """1M triangles with the reference benchmark's branching factor 4
    (benchmark_test.go:22): the tree bit-exact, closest hits (the benchmark's
    persistent 128-thread traversal) and traversalSteps equal the oracle's."""
    from trace_amd.device import set_tune

    print("phr n-ary 1M bf=4", flush=True)
    pos = scenes.random_triangles(1_000_000, seed=1)
    _, aux, phr, otree, ophr = _compare(ctx, oracle, pos, branching=4)
    assert phr.cost() < otree.cost()
    inc = scenes.incoherent_rays(200_000, seed=2)
    coh = scenes.coherent_rays_cfg2(4096, 2048)[::41]
    set_tune(ctx, trav_threads=128, trav_persistent=True)
    try:
        for rays in (inc, coh):
            hit, t = phr.closest_hit(rays)
            ohit, ot = ophr.closest_hit(rays)
            assert np.mean(hit == ohit) >= 0.9999
            both = (hit == ohit) & (hit >= 0)
            assert np.max(np.abs(t[both] - ot[both]) / ot[both]) < 1e-5
    finally:
        set_tune(ctx)
    r64 = inc[:20_000].astype(np.float64)
    assert np.array_equal(phr.traversal_steps(r64), ophr.traversal_steps(r64))


@pytest.mark.parametrize("bf", [3, 4, 8])
def test_phr_nary_steps_heatmap_render(ctx, oracle, bf):
    """traversalSteps and the heat map walk the n-ary tree's binary split but
    count only the tree's own nodes; a render over the n-ary tree matches the
    oracle's; the terrain's big leaves are split for traversal."""
    from trace_amd.device import render

    print(f"phr n-ary steps/heatmap/render bf={bf}", flush=True)
    pos = scenes.random_triangles(50_000, seed=51, mean_edge=0.01)
    lb = Bvh.lbvh(ctx, Mesh(ctx, pos))
    phr = Bvh.phr(ctx, lb, branching=bf)
    ref = oracle.PhrTree(oracle.Tree(pos), branching=bf)
    rays = scenes.incoherent_rays(20_000, seed=52).astype(np.float64)
    got = phr.traversal_steps(rays)
    want = ref.traversal_steps(rays)
    assert np.array_equal(got, want)
    assert want.max() > 10
    sc = scenes.box_scene(seed=1, slices=40, rings=22, width=160, height=90)
    m2 = Mesh(ctx, sc["pos"])
    m2.set_shading(sc["mat_idx"], sc["materials"])
    p2 = Bvh.phr(ctx, Bvh.lbvh(ctx, m2), branching=bf)
    ot2 = oracle.Tree(sc["pos"])
    rgb, counts = p2.heatmap(sc["camera"], 160, 90, 20)
    orgb, ocounts = oracle.PhrTree(ot2, branching=bf).heatmap(sc["camera"], 160, 90, 20)
    assert np.array_equal(counts, ocounts) and np.array_equal(rgb, orgb)
    s, q, _ = render(ctx, m2, p2, sc["camera"], 160, 90, 8, 5, _lib.MISS_SKY, seed=4)
    os_, oq, _ = oracle.render(ot2, sc["mat_idx"], sc["materials"], sc["camera"], 160, 90, 8, 5, _lib.MISS_SKY,
                               seed=4)
    from test_render_gpu import within_3_sigma

    frac, _ = within_3_sigma(s, q, os_, oq, 8)
    assert frac < 0.01
    d = scenes.terrain_scene(n=256, n_spheres=100, width=128, height=72)
    _, _, tphr, tot, _ = _compare(ctx, oracle, d["pos"], branching=bf)
    lo, hi = d["pos"].reshape(-1, 3).min(0), d["pos"].reshape(-1, 3).max(0)
    inc = scenes.incoherent_rays(20_000, seed=31)
    inc[:, :3] = lo + inc[:, :3] * (hi - lo)
    hit, t = tphr.closest_hit(inc)
    ohit, _ = tot.closest_hit(inc)
    assert np.mean(hit == ohit) >= 0.9999
