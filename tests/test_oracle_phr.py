"""Consistency checks of the oracle's PHR refinement (CPU only).

The oracle's PHR is pinned by what the mathematics fixes in
test_oracle_phr_pins.py (delta = 0 gives the LBVH, every SweepSAH call equals
a brute-force argmin, the refinement chain, the initial cut's breadth-first
characterisation, threshold equalities and hand-computed trees).  Here a
second, pure-Python restatement of PhrBuilder.BuildFromAuxilary
(phr.go:47-459), run on small inputs, must also give the oracle's tree
exactly: preorder topology, leaf triangle lists and FP64 boxes.
It works on the oracle's LBVH (itself pinned in test_oracle.py) and takes the
initial cut in FIFO order from the root, the order the oracle defines (one
worker draining the reference's channel queue).  Cost() of the PHR tree is
recomputed from the exported tree with bvh.go:226-239; invariants (every
triangle in exactly one leaf, parents contain children, at most two
children) are checked on a larger tree."""
import math
from collections import deque

import numpy as np
import pytest

from trace_amd import scenes

MAX_CUT = 2500


class N:
    """A BVH node: box (lo, hi tuples), children, prims (leaves)."""

    __slots__ = ("lo", "hi", "kids", "prims", "size", "aux")

    def __init__(self, lo, hi, kids=None, prims=None, aux=False):
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.kids = kids or []
        self.prims = prims
        self.size = 0
        self.aux = aux

    @property
    def leaf(self):
        return self.prims is not None

    def surface(self):  # bounding.go:45-47
        w, h, d = (self.hi[0] - self.lo[0]), (self.hi[1] - self.lo[1]), (self.hi[2] - self.lo[2])
        return 2 * w * h + 2 * w * d + 2 * h * d

    def bary(self, a):  # aabb.update, bounding.go:36-43
        return (self.lo[a] + self.hi[a]) * (1.0 / 2.0)

    def subtree_size(self):  # bvh.go:159-167
        if self.size == 0:
            self.size = 1 + sum(k.subtree_size() for k in self.kids)
        return self.size

    def collect_leaves(self, acc):
        if self.leaf:
            acc.append(self)
        else:
            for k in self.kids:
                k.collect_leaves(acc)


def union(a, b):  # aabb.add with MinVec/MaxVec
    lo = tuple(b[0][i] if a[0][i] > b[0][i] else a[0][i] for i in range(3))
    hi = tuple(b[1][i] if a[1][i] < b[1][i] else a[1][i] for i in range(3))
    return lo, hi


def box_surface(b):
    return N(b[0], b[1]).surface()


def lbvh_from_oracle(tree):
    """Rebuild the oracle LBVH from its preorder export."""
    rec, box = tree.canonical()
    it = iter(range(len(rec)))

    def build():
        i = next(it)
        kind, first, last, nch = (int(v) for v in rec[i])
        lo, hi = box[i, :3], box[i, 3:]
        if kind == 1:
            return N(lo, hi, prims=list(range(first, last + 1)), aux=True)
        kids = [build() for _ in range(nch)]
        return N(lo, hi, kids=kids, aux=True)

    return build()


class Phr:
    def __init__(self, alpha, delta, bf, max_cut):
        self.alpha, self.delta, self.bf, self.max_cut = alpha, delta, bf, max_cut

    def threshold(self, depth):  # DefaultThreshold, phr.go:356-358
        return self.surface / math.pow(2, self.alpha * float(depth) + float(self.delta))

    def initial_cut(self, root):  # phr.go:257-311, FIFO order, one worker
        self.cut_count = 0
        q = deque([root])
        cut = []
        while q:
            node = q.popleft()
            if node.leaf:
                cut.append(node)
            elif node.surface() > self.threshold(0):
                self.cut_count += len(node.kids) - 1
                if self.cut_count >= self.max_cut:
                    cut.append(node)
                    continue
                q.extend(node.kids)
            else:
                cut.append(node)
        return {"nodes": cut, "box": (root.lo, root.hi), "depth": 1}

    def refined(self, cut):  # phr.go:313-338
        out = []
        for node in cut["nodes"]:
            if node.leaf or node.surface() < self.threshold(cut["depth"]):
                out.append(node)
            else:
                out.extend(node.kids)
        if len(out) == 1:
            out[0] = make_leaf(cut["box"], out)
        return {"nodes": out, "box": cut["box"], "depth": cut["depth"] + 1}

    def build(self, aux_root):
        self.surface = aux_root.surface()
        self.split_work = 0  # buildSubTreeCost's counter (phr.go:195-255)
        return self.sub_tree(self.initial_cut(aux_root))

    def sub_tree(self, cut):  # buildSubTree, phr.go:92-152
        if len(cut["nodes"]) <= 1:
            return cut["nodes"][0]
        cuts = [cut]
        while len(cuts) < self.bf:
            mx, mi = 0, 0
            for i, c in enumerate(cuts):
                if len(c["nodes"]) > mx:
                    mx, mi = len(c["nodes"]), i
            if mx <= 1:
                break
            self.split_work += mx
            left, right = sweep_sah(cuts[mi])
            if right is not None:
                cuts[mi] = self.refined(left)
                cuts.append(self.refined(right))
            else:
                cuts[mi] = {"nodes": [make_leaf(left["box"], left["nodes"])], "box": left["box"],
                            "depth": left["depth"]}
        branch = N(cut["box"][0], cut["box"][1])
        branch.kids = [self.sub_tree(c) for c in cuts]
        return branch


def make_leaf(box, nodes):  # phr.go:340-352
    leaves = []
    for n in nodes:
        n.collect_leaves(leaves)
    prims = [p for leaf in leaves for p in leaf.prims]
    return N(box[0], box[1], prims=prims)


def min_cost(nodes, box, depth):  # phr.go:414-459
    n = len(nodes)
    sa_right = (nodes[-1].lo, nodes[-1].hi)
    right_costs = [0.0] * n
    right_cuts = [None] * n
    count = 0
    for i in range(n - 1, 0, -1):
        sa_right = union(sa_right, (nodes[i].lo, nodes[i].hi))
        count += nodes[i].subtree_size()
        right_costs[i] = box_surface(sa_right) * float(count)
        right_cuts[i] = {"nodes": nodes[i:], "box": sa_right, "depth": depth}
    count += nodes[0].subtree_size()
    best = {"cost": box_surface(box) * float(count), "left": {"nodes": nodes, "box": box, "depth": depth},
            "right": None}
    count = nodes[0].subtree_size()
    sa_left = (nodes[0].lo, nodes[0].hi)
    for i in range(1, n):
        cost = right_costs[i] + box_surface(sa_left) * float(count)
        if cost < best["cost"]:
            best = {"cost": cost, "left": {"nodes": nodes[:i], "box": sa_left, "depth": depth},
                    "right": right_cuts[i]}
        sa_left = union(sa_left, (nodes[i].lo, nodes[i].hi))
        count += nodes[i].subtree_size()
    return best


def sweep_sah(cut):  # phr.go:368-405 (sorted() is stable, as sort.SliceStable)
    cut["nodes"] = sorted(cut["nodes"], key=lambda nd: nd.bary(0))
    sorted2 = sorted(cut["nodes"], key=lambda nd: nd.bary(1))
    x = min_cost(cut["nodes"], cut["box"], cut["depth"])
    y = min_cost(sorted2, cut["box"], cut["depth"])
    if x["cost"] < y["cost"]:
        sorted2 = sorted(sorted2, key=lambda nd: nd.bary(2))
        z = min_cost(sorted2, cut["box"], cut["depth"])
        best = x if x["cost"] < z["cost"] else z
    else:
        cut["nodes"] = sorted(cut["nodes"], key=lambda nd: nd.bary(2))
        z = min_cost(cut["nodes"], cut["box"], cut["depth"])
        best = y if y["cost"] < z["cost"] else z
    return best["left"], best["right"]


def canonical(root):
    """Same format as the oracle: rec (kind, prim count or -1, n_children, 0),
    FP64 boxes, leaf prims in preorder."""
    rec, box, prims = [], [], []

    def walk(nd):
        if nd.leaf:
            rec.append((1, len(nd.prims), 0, 0))
            prims.extend(nd.prims)
        else:
            rec.append((0, -1, len(nd.kids), 0))
        box.append(nd.lo + nd.hi)
        for k in nd.kids:
            walk(k)

    walk(root)
    return np.array(rec, np.int32), np.array(box, np.float64), np.array(prims, np.int32)


def cost(nd):  # bvh.go:226-239, INTERSECTION_COST 1, TRAVERSAL_COST 2
    if nd.leaf:
        return 1.0 * float(len(nd.prims))
    s = 0.0
    for k in nd.kids:
        if nd.surface() != 0:
            s += (k.surface() / nd.surface()) * cost(k)
    return 2.0 + s


CASES = [
    (2, 1, 0.5, 6.0, MAX_CUT), (5, 2, 0.5, 6.0, MAX_CUT), (40, 3, 0.5, 6.0, MAX_CUT),
    (150, 4, 0.5, 6.0, MAX_CUT), (300, 5, 0.55, 9.0, MAX_CUT), (300, 6, 0.5, 6.0, 7),
    (200, 7, 0.0, 2.0, MAX_CUT), (200, 8, 1.0, 0.0, MAX_CUT), (120, 9, 0.5, 12.0, 3),
]


@pytest.mark.parametrize("n,seed,alpha,delta,max_cut", CASES)
def test_phr_oracle_equals_python_restatement(oracle, n, seed, alpha, delta, max_cut):
    pos = scenes.random_triangles(n, seed=seed, mean_edge=0.05)
    aux = oracle.Tree(pos)
    want = oracle.PhrTree(aux, alpha, delta, 2, max_cut)
    py = Phr(alpha, delta, 2, max_cut)
    root = py.build(lbvh_from_oracle(aux))
    rec, box, prims = canonical(root)
    orec, obox, oprims = want.canonical()
    assert rec.shape == orec.shape, (rec.shape, orec.shape)
    assert np.array_equal(rec, orec)
    assert np.array_equal(box, obox)
    assert np.array_equal(prims, oprims)
    assert want.cost() == cost(root)
    n_cut = len(py.initial_cut(lbvh_from_oracle(aux))["nodes"])
    assert want.initial_cut() == n_cut
    # BuildWithCost's cost: the initial cut plus every cut handed to Split
    assert want.build_cost() == n_cut + py.split_work


def test_phr_duplicates_and_flat(oracle):
    base = scenes.random_triangles(60, seed=12, mean_edge=0.05)
    pos = np.concatenate([base, base[:20], base[:5]])
    pos.reshape(-1, 3, 3)[:30, :, 1] = 0.5
    aux = oracle.Tree(pos)
    root = Phr(0.5, 6.0, 2, MAX_CUT).build(lbvh_from_oracle(aux))
    rec, box, prims = canonical(root)
    orec, obox, oprims = oracle.PhrTree(aux).canonical()
    assert np.array_equal(rec, orec) and np.array_equal(box, obox) and np.array_equal(prims, oprims)


def test_phr_invariants_larger_tree(oracle):
    pos = scenes.random_triangles(20_000, seed=13)
    aux = oracle.Tree(pos)
    phr = oracle.PhrTree(aux)
    rec, box, prims = phr.canonical()
    assert np.array_equal(np.sort(prims), np.arange(len(pos)))  # every triangle in exactly one leaf
    assert np.all(rec[:, 2] <= 2)
    # parents contain children: walk the preorder with a stack of open parents
    stack = []
    for i in range(len(rec)):
        if stack:
            p = stack[-1]
            assert np.all(box[p, :3] <= box[i, :3]) and np.all(box[i, 3:] <= box[p, 3:])
            remaining[p] -= 1
            if remaining[p] == 0:
                stack.pop()
        if rec[i, 0] == 0 and rec[i, 2] > 0:
            if not stack:
                remaining = {}
            remaining[i] = int(rec[i, 2])
            stack.append(i)
        while stack and remaining[stack[-1]] == 0:
            stack.pop()
    # leaf boxes contain their triangles
    _, perm = aux.sorted()
    off = 0
    tri = pos.reshape(-1, 3, 3)
    for i in np.nonzero(rec[:, 0] == 1)[0]:
        cnt = int(rec[i, 1])
        t = tri[perm[prims[off:off + cnt]]]
        assert np.all(box[i, :3] <= t.min(axis=(0, 1))) and np.all(t.max(axis=(0, 1)) <= box[i, 3:])
        off += cnt
    assert phr.cost() < aux.cost()  # refinement lowers the SAH cost of the LBVH


def test_grid_optimizer_selection(oracle):
    """optimizer.go:44-94: with omega ~ 0 the score is the build cost alone,
    with a huge omega the SAH term decides; the first minimum wins."""
    from oracle.pyoracle import grid_optimize

    pos = scenes.random_triangles(3000, seed=21, mean_edge=0.02)
    tree = oracle.Tree(pos)
    alphas, deltas = (0.4, 0.5, 0.55), (5.0, 6.0, 9.0)
    a, d, ev = grid_optimize(tree, alphas, deltas, 2, 2)  # omega ~ 5e-10
    sah = np.array([e[0] for e in ev])
    build = np.array([e[1] for e in ev])
    assert len(set(build.tolist())) > 1 and len(set(sah.tolist())) > 1
    k = int(np.argmin(build))
    assert (a, d) == (alphas[k // 3], deltas[k % 3])
    a, d, _ = grid_optimize(tree, alphas, deltas, 3_000_000, 3_000_000)  # omega ~ 100
    k = int(np.argmin(sah))
    assert (a, d) == (alphas[k // 3], deltas[k % 3])
    # every evaluation is BuildWithCost + Cost of that (alpha, delta)
    p = oracle.PhrTree(tree, alphas[1], deltas[2])
    assert ev[5] == (p.cost(), p.build_cost())


def test_omega_and_eval_phr_closed_forms():
    """optimizer.go:88-94 at values whose results are exact: a dropped square,
    a swapped ratio or a lost term changes each of them."""
    from oracle.pyoracle import eval_phr, omega

    assert omega(2_000_000, 200_000) == 1.0
    assert omega(1_000_000, 400_000) == 0.5  # (1/2)^2 * 2
    assert omega(4_000_000, 100_000) == 2.0  # 2^2 * (1/2)
    assert eval_phr(1.0, 4.0, 3, 6, 2.0) == 1.5  # 3/6 + 2^2 * 1/4
    assert eval_phr(5.0, 5.0, 7, 7, 0.0) == 1.0
    assert eval_phr(2.0, 8.0, 0, 9, 3.0) == 2.25  # 0 + 9 * 1/4


def test_bayes_objective_closed_forms(oracle):
    """optimizer.go:132-151: at the range minimum the SAH ratio is 1, at the
    range maximum the build ratio is 1, and a zero parameter scores the two
    normalisers, 1 + omega^2."""
    from oracle.pyoracle import bayes_objective, omega

    pos = scenes.random_triangles(3000, seed=22, mean_edge=0.02)
    tree = oracle.Tree(pos)
    ar, dr = (0.4, 0.55), (6.0, 9.0)
    om = omega(tree.n, 640 * 480)
    lo = oracle.PhrTree(tree, ar[0], dr[0])
    hi = oracle.PhrTree(tree, ar[1], dr[1])
    y, sah, build = bayes_objective(tree, ar[0], dr[0], ar, dr, 640, 480)
    assert (sah, build) == (lo.cost(), lo.build_cost())
    assert y == float(build) / float(hi.build_cost()) + om * om
    y, sah, build = bayes_objective(tree, ar[1], dr[1], ar, dr, 640, 480)
    assert y == 1.0 + (om * om) * (hi.cost() / lo.cost())
    y, sah, build = bayes_objective(tree, 0.0, 7.0, ar, dr, 640, 480)
    assert (sah, build) == (lo.cost(), hi.build_cost()) and y == 1.0 + om * om
    # a point between the ends is that point's BuildWithCost + Cost
    p = oracle.PhrTree(tree, 0.47, 7.5)
    y, sah, build = bayes_objective(tree, 0.47, 7.5, ar, dr, 640, 480)
    assert (sah, build) == (p.cost(), p.build_cost())
