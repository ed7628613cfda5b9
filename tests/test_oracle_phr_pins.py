"""Pins of the oracle's PHR refinement (phr.go) by what the mathematics fixes
(CPU only).

- delta = 0: the root's surface is not above S / 2^0, so the initial cut is the
  root alone and the PHR tree is the LBVH itself (phr.go:287, 93-96).
- Every SweepSAH call (phr.go:368-459) the oracle makes is recomputed by brute
  force from the members alone: each member's box is the union of its
  triangles' boxes, its weight the size of its LBVH subtree (2 U - 1 nodes
  over U distinct Morton keys, bvh.go:159-167); for each axis (the stable
  sort chain x, y-of-x, z-of-the-loser) every split position is costed
  directly as SA(left) * sum(left) + SA(right) * sum(right) against the
  no-split SA(bounding) * sum(all), first minimum winning; the oracle's left
  and right member lists must be exactly the argmin's.  Weighting by
  triangle count instead picks differently on the same corpus, so the check
  has the power to tell them apart.
- The job splits its largest cut first, the first of equal sizes (phr.go:102-115).
- The initial cut is the set of nodes whose every ancestor's surface exceeds
  S / 2^delta (leaves, or nodes at or below the threshold), in breadth-first
  order, and with the 2500 cap exactly cap - 1 nodes are expanded.
- Surfaces exactly equal to the threshold: the initial cut keeps such a node
  (">" expands, phr.go:287) while refinement expands it ("<" keeps, phr.go:319).
- A four-triangle scene whose tree, Cost and BuildWithCost are computed by hand."""
import numpy as np
import pytest

from trace_amd import scenes


# ---- independent views of the oracle's LBVH -------------------------------------


def tri_boxes(pos, perm):
    """FP64 boxes of the triangles in sorted order (Min3/Max3 of the vertices)."""
    v = pos[perm].reshape(-1, 3, 3)
    return np.concatenate([v.min(axis=1), v.max(axis=1)], axis=1)


def member_box(boxes, first, last):
    b = boxes[first:last + 1]
    return np.concatenate([b[:, :3].min(axis=0), b[:, 3:].max(axis=0)])


def subtree_size(keys, first, last):
    """Nodes of the LBVH subtree over sorted positions [first, last]: a binary
    radix tree over U distinct keys has U leaves and U - 1 inner nodes."""
    u = len(np.unique(keys[first:last + 1]))
    return 2 * u - 1


def surface(b):  # bounding.go:45-47, same operation order
    w, h, d = b[3] - b[0], b[4] - b[1], b[5] - b[2]
    return 2 * w * h + 2 * w * d + 2 * h * d


def union(boxes):
    return np.concatenate([boxes[:, :3].min(axis=0), boxes[:, 3:].max(axis=0)])


def axis_best(order, boxes, weights, bounding):
    """Smallest cost over every split position of `order`, costed directly;
    starts from the no-split cost and keeps the first strict minimum."""
    best_cost = surface(bounding) * float(weights.sum())
    best_i = 0
    for i in range(1, len(order)):
        L, R = order[:i], order[i:]
        c = surface(union(boxes[R])) * float(weights[R].sum()) + surface(union(boxes[L])) * float(weights[L].sum())
        if c < best_cost:
            best_cost, best_i = c, i
    return best_cost, best_i


def stable_by(order, centres, axis):
    return order[np.argsort(centres[order, axis], kind="stable")]


def brute_force_split(boxes, weights, bounding):
    """(left, right) member lists SweepSAH must return, or None for no split."""
    centres = (boxes[:, :3] + boxes[:, 3:]) * (1.0 / 2.0)
    m = len(boxes)
    xs = stable_by(np.arange(m), centres, 0)
    ys = stable_by(xs, centres, 1)
    x = axis_best(xs, boxes, weights, bounding)
    y = axis_best(ys, boxes, weights, bounding)
    if x[0] < y[0]:
        zs = stable_by(ys, centres, 2)
        z = axis_best(zs, boxes, weights, bounding)
        order, (c, i) = (xs, x) if x[0] < z[0] else (zs, z)
    else:
        zs = stable_by(xs, centres, 2)
        z = axis_best(zs, boxes, weights, bounding)
        order, (c, i) = (ys, y) if y[0] < z[0] else (zs, z)
    if i == 0:
        return None
    return order[:i], order[i:]


def refine(members, depth, S, alpha, delta, nodes, by_range, keys):
    """phr.go:313-338 on LBVH members: a member at or above the threshold of
    this depth is replaced by its two children; leaves stay."""
    th = S / 2.0 ** (alpha * float(depth) + float(delta))
    out = []
    for f, l in members:
        i = by_range[(f, l)]
        leaf = len(np.unique(keys[f:l + 1])) == 1
        if leaf or surface(nodes[i][5]) < th:
            out.append((f, l))
        else:
            out.extend(nodes[c][:2] for c in nodes[i][6])
    return out


def check_chain(tree, calls, alpha, delta):
    """Each call's members are the refinement of an earlier call's left or
    right side (or, for the first, the initial cut), each used once; every
    refined side of two or more members is split later."""
    keys, _ = tree.sorted()
    nodes = lbvh_nodes(tree)
    by_range = {tuple(nd[:2]): i for i, nd in enumerate(nodes)}
    S = surface(nodes[0][5])
    pending = []  # (members, depth) of cuts still to be split
    for k, call in enumerate(calls):
        members = [tuple(r) for r in call["ranges"]]
        if k == 0:
            depth = 1
        else:
            match = [j for j, (m, _d) in enumerate(pending) if m == members]
            assert match, f"call {k}: members are no refined side of an earlier split"
            depth = pending.pop(match[0])[1]
        if call["split"]:
            for side in (call["left"], call["right"]):
                ref = refine([members[i] for i in side], depth, S, alpha, delta, nodes, by_range, keys)
                if len(ref) > 1:
                    pending.append((ref, depth + 1))
    assert not pending, f"{len(pending)} refined cuts were never split"


def check_calls(tree, pos, calls):
    """Every traced SweepSAH call against the brute force; returns how many
    calls the triangle-count weighting would have split differently."""
    keys, perm = tree.sorted()
    boxes = tri_boxes(pos, perm)
    differs = 0
    for k, call in enumerate(calls):
        sizes = call["cut_sizes"]
        # the largest cut is split, the first of equal sizes
        assert call["chosen"] == int(np.argmax(sizes)), (k, sizes)
        assert sizes[call["chosen"]] == len(call["ranges"]) > 1
        rng = call["ranges"]
        assert np.all(rng[:, 0] >= 0), "members are LBVH nodes"
        mb = np.array([member_box(boxes, f, l) for f, l in rng])
        assert np.array_equal(mb, call["boxes"]), k
        # a cut's bounding is the union of its members' boxes
        assert np.array_equal(call["bounding"], union(mb)), k
        w = np.array([subtree_size(keys, f, l) for f, l in rng], np.int64)
        want = brute_force_split(mb, w, call["bounding"])
        if want is None:
            assert not call["split"], k
        else:
            assert call["split"], k
            assert np.array_equal(call["left"], want[0]), (k, call["left"], want[0])
            assert np.array_equal(call["right"], want[1]), (k, call["right"], want[1])
        prim_w = rng[:, 1] - rng[:, 0] + 1
        other = brute_force_split(mb, prim_w.astype(np.int64), call["bounding"])
        if (other is None) != (want is None) or (other is not None and not (
                np.array_equal(other[0], want[0]) and np.array_equal(other[1], want[1]))):
            differs += 1
    return differs


# ---- delta = 0 --------------------------------------------------------------------


@pytest.mark.parametrize("n,seed,alpha", [(1, 1, 0.5), (2, 2, 0.0), (57, 3, 0.5), (1000, 4, 2.0)])
def test_delta_zero_gives_the_lbvh(oracle, n, seed, alpha):
    pos = scenes.random_triangles(n, seed=seed, mean_edge=0.05)
    aux = oracle.Tree(pos)
    phr = oracle.PhrTree(aux, alpha, 0.0, 2)
    rec, box, prims = phr.canonical()
    arec, abox = aux.canonical()
    assert phr.initial_cut() == 1
    assert phr.build_cost() == 1  # the initial cut; a one-node cut is attached, never split
    assert np.array_equal(rec[:, 0], arec[:, 0])
    assert np.array_equal(rec[:, 2], arec[:, 3])
    leaf = arec[:, 0] == 1
    assert np.array_equal(rec[leaf, 1], arec[leaf, 2] - arec[leaf, 1] + 1)
    assert np.array_equal(prims, np.concatenate([np.arange(f, l + 1) for f, l in arec[leaf, 1:3]]))
    assert np.array_equal(box, abox)
    assert phr.cost() == aux.cost()


# ---- every split is the brute-force argmin -------------------------------------------


SPLIT_CASES = [
    # n, seed, alpha, delta, max_cut
    (3, 11, 0.5, 6.0, 2500), (5, 12, 0.0, 9.0, 2500), (8, 13, 0.5, 1.0, 2500), (8, 14, 0.0, 30.0, 2500),
    (12, 15, 0.3, 2.0, 2500), (30, 16, 0.5, 3.0, 2500), (40, 17, 1.0, 2.0, 2500), (60, 18, 0.5, 6.0, 2500),
    (80, 19, 0.0, 4.0, 2500), (120, 20, 0.55, 9.0, 2500), (150, 21, 0.5, 6.0, 9), (200, 22, 0.25, 5.0, 2500),
]


@pytest.mark.parametrize("bf", [2, 3, 4])
def test_every_split_is_the_brute_force_argmin(oracle, bf):
    differs, calls_total = 0, 0
    for n, seed, alpha, delta, cap in SPLIT_CASES:
        pos = scenes.random_triangles(n, seed=seed, mean_edge=0.08)
        aux = oracle.Tree(pos)
        phr = oracle.PhrTree(aux, alpha, delta, bf, cap, trace=True)
        calls = phr.split_calls()
        calls_total += len(calls)
        differs += check_calls(aux, pos, calls)
        check_chain(aux, calls, alpha, delta)
        # BuildWithCost = initial cut + the size of every cut handed to Split
        assert phr.build_cost() == phr.initial_cut() + sum(len(c["ranges"]) for c in calls)
        rec, _, prims = phr.canonical()
        assert np.all(rec[:, 2] <= bf) and np.array_equal(np.sort(prims), np.arange(n))
    assert calls_total > 200
    # the corpus separates subtree-size weighting from triangle-count weighting
    assert differs > 0


def layered(n, seed, spread):
    """Triangles whose boxes lie in two thin layers (z in [0, 0.1] or
    [5, 5.1]): their z centres tie within a layer, so the order a stable z
    sort starts from (y's order after x wins over y, x's after y wins)
    decides the order of the members it returns.  `spread` scales the x and
    y ranges, so either axis can be the one that loses."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x0, y0 = rng.random(2) * np.array(spread)
        w, h = 0.05 + 0.2 * rng.random(2)
        z = 5.0 if rng.random() < 0.5 else 0.0
        out.append([x0, y0, z, x0 + w, y0 + h, z, x0, y0 + h, z + 0.1])
    return np.array(out, np.float32).astype(np.float64)


@pytest.mark.parametrize("seed,spread", [(s, sp) for s in range(4) for sp in ((1.0, 3.0), (3.0, 1.0))])
def test_splits_with_tied_centres(oracle, seed, spread):
    pos = layered(24, 61 + seed, spread)
    aux = oracle.Tree(pos)
    for alpha, delta, bf in ((0.0, 12.0, 2), (0.5, 2.0, 2), (0.0, 12.0, 4)):
        phr = oracle.PhrTree(aux, alpha, delta, bf, trace=True)
        calls = phr.split_calls()
        check_calls(aux, pos, calls)
        check_chain(aux, calls, alpha, delta)


def test_splits_with_duplicate_keys(oracle):
    """Equal Morton keys collapse into one LBVH leaf: such a member weighs 1
    (one node) however many triangles it holds."""
    base = scenes.random_triangles(40, seed=31, mean_edge=0.08)
    pos = np.concatenate([base, base[:15], base[:15], base[3:7]])
    aux = oracle.Tree(pos)
    phr = oracle.PhrTree(aux, 0.2, 3.0, 2, trace=True)
    calls = phr.split_calls()
    assert calls
    check_calls(aux, pos, calls)
    check_chain(aux, calls, 0.2, 3.0)
    assert any(np.any(c["ranges"][:, 1] > c["ranges"][:, 0]) for c in calls)


# ---- the initial cut ------------------------------------------------------------------


def lbvh_nodes(tree):
    """Preorder nodes of the oracle LBVH: (first, last, depth, parent, is_leaf,
    box, children)."""
    rec, box = tree.canonical()
    nodes, stack = [], []
    for i, (kind, first, last, nch) in enumerate(rec):
        parent = stack[-1][0] if stack else -1
        depth = len(stack)
        nodes.append((int(first), int(last), depth, parent, kind == 1, box[i], []))
        if parent >= 0:
            nodes[parent][6].append(i)
        if stack:
            stack[-1][1] -= 1
        if kind == 0:
            stack.append([i, int(nch)])
        while stack and stack[-1][1] == 0:
            stack.pop()
    return nodes


@pytest.mark.parametrize("n,seed,delta,cap", [(300, 41, 6.0, 2500), (500, 42, 9.0, 2500), (400, 43, 12.0, 25),
                                              (200, 44, 20.0, 2)])
def test_initial_cut_is_breadth_first_below_the_threshold(oracle, n, seed, delta, cap):
    pos = scenes.random_triangles(n, seed=seed, mean_edge=0.05)
    aux = oracle.Tree(pos)
    nodes = lbvh_nodes(aux)
    S = surface(nodes[0][5])
    th = S / 2.0 ** delta  # threshold at depth 0 (alpha * 0 + delta)
    phr = oracle.PhrTree(aux, 0.5, delta, 2, cap)
    got = [tuple(r) for r in phr.initial_ranges()]
    by_range = {(f, l): i for i, (f, l, *_rest) in enumerate(nodes)}
    ids = [by_range[r] for r in got]
    # disjoint and covering every sorted position
    cover = np.concatenate([np.arange(f, l + 1) for f, l in got])
    assert np.array_equal(np.sort(cover), np.arange(n))
    # breadth-first order: depth never decreases, left to right within a depth
    keyed = [(nodes[i][2], nodes[i][0]) for i in ids]
    assert keyed == sorted(keyed)
    # nodes above a member were expanded: their surface exceeds the threshold
    expanded = set()
    for i in ids:
        p = nodes[i][3]
        while p >= 0:
            assert surface(nodes[p][5]) > th
            expanded.add(p)
            p = nodes[p][3]
    expandable = [i for i in ids if not nodes[i][4] and surface(nodes[i][5]) > th]
    if expandable:  # members kept only because the cap was reached
        assert len(expanded) == cap - 1
    else:
        assert len(expanded) < cap


# ---- surfaces exactly at the threshold ------------------------------------------------


def octant_pairs():
    """16 triangles, two per octant of [0,1]^3: their boxes are the lower and
    upper z-halves of the octant, so every LBVH box has dyadic coordinates
    (root surface 6, z-halves 4, quarters 2.5, octants 1.5, triangles 1)."""
    tris = []
    for oz in (0.0, 0.5):
        for oy in (0.0, 0.5):
            for ox in (0.0, 0.5):
                for dz in (0.0, 0.25):
                    z0 = oz + dz
                    tris.append([ox, oy, z0, ox + 0.5, oy + 0.5, z0, ox, oy + 0.5, z0 + 0.25])
    return np.array(tris, np.float64)


def test_threshold_equality_kept_by_initial_cut_expanded_by_refine(oracle):
    pos = octant_pairs()
    aux = oracle.Tree(pos)
    nodes = lbvh_nodes(aux)
    assert surface(nodes[0][5]) == 6.0
    # alpha 0, delta 2: the threshold is 6 / 4 = 1.5 at every depth, an octant's surface
    phr = oracle.PhrTree(aux, 0.0, 2.0, 2, trace=True)
    cut = phr.initial_ranges()
    assert len(cut) == 8 and np.all(cut[:, 1] - cut[:, 0] == 1)  # the octants, kept whole
    calls = phr.split_calls()
    assert len(calls[0]["ranges"]) == 8
    # after the first split each side is refined: octants (1.5, not below 1.5)
    # are replaced by their two triangles, so later calls see only leaves
    for c in calls[1:]:
        assert np.all(c["ranges"][:, 0] == c["ranges"][:, 1])
    assert len(calls[1]["ranges"]) == 8
    check_calls(aux, pos, calls)
    check_chain(aux, calls, 0.0, 2.0)
    rec, box, prims = phr.canonical()
    assert np.array_equal(np.sort(prims), np.arange(16))
    assert np.all(rec[rec[:, 0] == 1, 1] == 1)  # 16 one-triangle leaves


# ---- a tree computed by hand ----------------------------------------------------------


def test_four_cubes_by_hand(oracle):
    """Four triangles whose boxes are unit cubes side by side along x.  LBVH:
    {T0, T1}, {T2, T3} (x's top Morton bit).  alpha 0, delta 1: threshold
    18 / 2 = 9; the root (surface 18) and both pairs (10) are expanded, so the
    initial cut is T0..T3 (BuildWithCost 4).  Root job: no split costs
    18 * 4 = 72; along x the splits cost 6 + 42, 20 + 20, 42 + 6; y and z keep
    x's order and the same costs, so the tie goes to z: {T0, T1} | {T2, T3}
    (+4).  Each pair splits at 6 + 6 < 20 (+2, +2) into one-triangle leaves
    whose boxes are the triangles'.  BuildWithCost = 12; Cost = 2 + 2 (10/18)
    (2 + 2 (6/10))."""
    pos = np.array([[x, 0, 0, x + 1, 1, 0, x, 1, 1] for x in range(4)], np.float64)
    aux = oracle.Tree(pos)
    keys, perm = aux.sorted()
    assert np.array_equal(perm, np.arange(4))
    phr = oracle.PhrTree(aux, 0.0, 1.0, 2)
    assert phr.initial_cut() == 4
    assert phr.build_cost() == 12
    rec, box, prims = phr.canonical()
    want_rec = np.array([[0, -1, 2, 0], [0, -1, 2, 0], [1, 1, 0, 0], [1, 1, 0, 0],
                         [0, -1, 2, 0], [1, 1, 0, 0], [1, 1, 0, 0]], np.int32)
    assert np.array_equal(rec, want_rec)
    cube = lambda x0, x1: [x0, 0, 0, x1, 1, 1]  # noqa: E731
    want_box = np.array([cube(0, 4), cube(0, 2), cube(0, 1), cube(1, 2), cube(2, 4), cube(2, 3), cube(3, 4)],
                        np.float64)
    assert np.array_equal(box, want_box)
    assert np.array_equal(prims, np.arange(4))
    assert phr.cost() == pytest.approx(2 + 2 * (10 / 18) * (2 + 2 * (6 / 10)), rel=1e-15)


def test_three_cubes_by_hand(oracle):
    """Three unit cubes along x: the LBVH is ((T0, T1), T2) (x's top Morton
    bit: 0, 0, 1).  alpha 0, delta 3: threshold 14 / 8; every inner node is
    expanded, and the breadth-first initial cut is T2, T0, T1 (T2, a leaf
    child of the root, is dequeued before T0 and T1).  The root job sorts them
    to T0, T1, T2 on every axis; both splits cost 6 * 1 + 10 * 2 = 26 < 14 * 3,
    an exact tie that the first minimum settles: T0 | T1, T2.  {T1, T2} splits
    at 6 + 6 < 20.  The PHR tree (T0, (T1, T2)) is not the LBVH;
    BuildWithCost = 3 + 3 + 2; Cost = 2 + (6/14) 1 + (10/14)(2 + 2 (6/10))."""
    pos = np.array([[x, 0, 0, x + 1, 1, 0, x, 1, 1] for x in range(3)], np.float64)
    aux = oracle.Tree(pos)
    arec, _ = aux.canonical()
    assert np.array_equal(arec[:, 0], [0, 0, 1, 1, 1])  # ((T0, T1), T2)
    phr = oracle.PhrTree(aux, 0.0, 3.0, 2)
    assert np.array_equal(phr.initial_ranges(), [[2, 2], [0, 0], [1, 1]])
    assert phr.build_cost() == 8
    rec, box, prims = phr.canonical()
    want_rec = np.array([[0, -1, 2, 0], [1, 1, 0, 0], [0, -1, 2, 0], [1, 1, 0, 0], [1, 1, 0, 0]], np.int32)
    assert np.array_equal(rec, want_rec)
    cube = lambda x0, x1: [x0, 0, 0, x1, 1, 1]  # noqa: E731
    assert np.array_equal(box, np.array([cube(0, 3), cube(0, 1), cube(1, 3), cube(1, 2), cube(2, 3)], np.float64))
    assert np.array_equal(prims, [0, 1, 2])
    assert phr.cost() == pytest.approx(2 + 6 / 14 + (10 / 14) * (2 + 2 * (6 / 10)), rel=1e-15)
