"""Pins the FP64 oracle (CPU only) against checks that do not restate it:
naive bit loops, closed forms, brute force and a by-definition radix tree.
Each check is chosen so that a plausible slip in the oracle (a swapped axis,
a wrong shift, an off-by-one split, a dropped tie rule) fails it."""
import math

import numpy as np
import pytest

from trace_amd import scenes

REF_DIR_CUBE = None


def naive_interleave(x: int, y: int, z: int) -> int:
    out = 0
    for i in range(21):
        out |= ((x >> i) & 1) << (3 * i)
        out |= ((y >> i) & 1) << (3 * i + 1)
        out |= ((z >> i) & 1) << (3 * i + 2)
    return out


def test_split_by3_equals_bit_loop(oracle):
    rng = np.random.default_rng(0)
    vals = [0, 1, 2, 3, (1 << 21) - 1, 1 << 20, 0x155555, 0x0AAAAA] + [int(v) for v in rng.integers(0, 1 << 21, 500)]
    for v in vals:
        assert oracle.split_by3(v) == naive_interleave(v, 0, 0)
        assert oracle.split_by3(v) << 1 == naive_interleave(0, v, 0)
    # bits above 21 are ignored (the reference masks with 0x1fffff)
    assert oracle.split_by3((1 << 21) | 5) == oracle.split_by3(5)


def _point_tri(p):
    return np.array([list(p) * 3], np.float64)


@pytest.mark.parametrize("box", [(0, 0, 0, 1, 1, 1), (-2.0, 0.5, -1.0, 3.0, 0.75, 11.5), (-1e-3, -2e-3, 0, 0, 0, 5e-3)])
def test_morton_closed_forms(oracle, box):
    box = np.array(box, np.float64)
    lo, hi = box[:3], box[3:]
    # a triangle collapsed onto the low corner quantises to 0 on every axis
    assert oracle.morton(_point_tri(lo), box)[0] == 0
    # ... onto the high corner to 2^21 - 1 on every axis: all 63 bits set
    assert int(oracle.morton(_point_tri(hi), box)[0]) == (1 << 63) - 1
    # M30 keeps the top 30 bits
    assert int(oracle.morton(_point_tri(hi), box, 1)[0]) == (1 << 30) - 1
    # only x at the top: x bits are the bit positions 0 mod 3
    p = lo.copy()
    p[0] = hi[0]
    assert int(oracle.morton(_point_tri(p), box)[0]) == naive_interleave((1 << 21) - 1, 0, 0)
    p = lo.copy()
    p[2] = hi[2]
    assert int(oracle.morton(_point_tri(p), box)[0]) == naive_interleave(0, 0, (1 << 21) - 1)


def test_morton_uses_box_centre_and_truncates(oracle):
    box = np.array([0, 0, 0, 1, 1, 1], np.float64)
    # box centre, not vertex mean: vertices (0,0,0),(0.5,0,0),(0.5,0,0) have
    # mean x = 1/3 but box centre x = 0.25
    tri = np.array([[0, 0, 0, 0.5, 0, 0, 0.5, 0, 0]], np.float64)
    qx = int(math.floor(0.25 / (1.0 / 2097151.0)))
    assert int(oracle.morton(tri, box)[0]) == naive_interleave(qx, 0, 0)
    # a zero-extent axis quantises to 0
    flat = np.array([0, 0, 0, 1, 1, 0], np.float64)
    tri = np.array([[1, 1, 0, 1, 1, 0, 1, 1, 0]], np.float64)
    assert int(oracle.morton(tri, flat)[0]) == naive_interleave((1 << 21) - 1, (1 << 21) - 1, 0)


def test_morton_matches_float_loop(oracle):
    # Python floats are IEEE binary64 with no contraction: a guard against the
    # oracle being compiled with FMA contraction or fast-math.
    pos = scenes.random_triangles(2000, seed=5)
    box = oracle.enclosing(pos)
    keys = oracle.morton(pos, box)
    for i in range(0, 2000, 7):
        q = []
        for a in range(3):
            vs = [pos[i, a], pos[i, 3 + a], pos[i, 6 + a]]
            c = (min(vs) + max(vs)) * 0.5
            f = abs(box[a] - c) / ((box[3 + a] - box[a]) / 2097151.0)
            q.append(int(f))
        assert int(keys[i]) == naive_interleave(*q)


def test_enclosing_is_bruteforce_union(oracle):
    pos = scenes.random_triangles(3000, seed=3)
    v = pos.reshape(-1, 3, 3)
    want = np.concatenate([v.min(axis=(0, 1)), v.max(axis=(0, 1))])
    assert np.array_equal(oracle.enclosing(pos), want)


def test_stable_sort(oracle):
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 50, 5000).astype(np.uint64)  # many ties
    s, perm = oracle.stable_sort(keys)
    assert np.all(np.diff(s.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(5000))
    assert np.array_equal(keys[perm], s)
    ties = s[1:] == s[:-1]
    assert np.all(perm[1:][ties] > perm[:-1][ties])  # equal keys keep index order


def radix_tree_by_definition(keys):
    """Preorder (kind, first, last, n_children) of the binary radix tree of a
    sorted key list: a run of equal keys is a leaf; otherwise split at the
    first key carrying the highest bit in which the first and last key differ."""
    out = []

    def rec(b, e):
        if keys[b] == keys[e]:
            out.append((1, b, e, 0))
            return
        bit = (int(keys[b]) ^ int(keys[e])).bit_length() - 1
        s = next(i for i in range(b, e + 1) if (int(keys[i]) >> bit) & 1)
        out.append((0, b, e, 2))
        rec(b, s - 1)
        rec(s, e)

    rec(0, len(keys) - 1)
    return np.array(out, np.int32)


@pytest.mark.parametrize("key_mode", [0, 1])
def test_lbvh_topology_by_definition(oracle, key_mode):
    pos = scenes.random_triangles(1500, seed=11)
    # duplicate some triangles so that equal keys occur
    pos = np.concatenate([pos, pos[:200], pos[100:150]])
    t = oracle.Tree(pos, key_mode=key_mode)
    keys, perm = t.sorted()
    rec, box = t.canonical()
    assert np.array_equal(rec, radix_tree_by_definition(keys))
    # leaf boxes are brute-force unions of their triangles' boxes
    v = pos.reshape(-1, 3, 3)
    for r, b in zip(rec, box):
        idx = perm[r[1]:r[2] + 1]
        lo = v[idx].min(axis=(0, 1))
        hi = v[idx].max(axis=(0, 1))
        assert np.array_equal(b, np.concatenate([lo, hi]))


def test_lbvh_single_triangle_is_root_leaf(oracle):
    t = oracle.Tree(scenes.random_triangles(1, seed=2))
    rec, _ = t.canonical()
    assert rec.tolist() == [[1, 0, 0, 0]]
    assert t.cost() == 1.0


def test_cost_hand_tree(oracle):
    # two unit triangles far apart: root box and leaf boxes computable by hand
    a = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    b = [9, 0, 0, 10, 0, 0, 9, 1, 0]
    t = oracle.Tree(np.array([a, b], np.float64))
    # leaves are flat (z extent 0): SA = 2*w*h = 2; root box 10 x 1 x 0: SA = 20
    assert t.cost() == pytest.approx(2.0 + 2 / 20 * 1 + 2 / 20 * 1, rel=0, abs=1e-15)


def _ray(o, d, tmin=0.001, tmax=math.inf):
    return np.array([[*o, tmin, *d, tmax]], np.float32)


def test_single_triangle_analytic(oracle):
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float64)
    t = oracle.Tree(tri)
    h, tt = t.closest_hit(_ray((0.25, 0.25, 2.0), (0, 0, -1)))
    assert h[0] == 0 and tt[0] == 2.0
    # unnormalised direction: t scales inversely
    h, tt = t.closest_hit(_ray((0.25, 0.25, 2.0), (0, 0, -4)))
    assert h[0] == 0 and tt[0] == 0.5
    # outside the triangle (u + v > 1)
    h, _ = t.closest_hit(_ray((0.75, 0.75, 2.0), (0, 0, -1)))
    assert h[0] == -1
    # t == tmax is accepted (inclusive), t > tmax is not
    h, _ = t.closest_hit(_ray((0.25, 0.25, 2.0), (0, 0, -1), tmax=2.0))
    assert h[0] == 0
    h, _ = t.closest_hit(_ray((0.25, 0.25, 2.0), (0, 0, -1), tmax=1.999))
    assert h[0] == -1
    # behind the origin / below tmin
    h, _ = t.closest_hit(_ray((0.25, 0.25, -2.0), (0, 0, -1)))
    assert h[0] == -1


def test_det_threshold_absolute(oracle):
    # det = e1 . (d x e2); with e1 = (s,0,0), e2 = (0,s,0), d = (0,0,-1): det = s^2
    for s, hits in [(1e-3, True), (4.0e-4, True), (3.0e-4, False), (1.0e-4, False)]:
        tri = np.array([[0, 0, 0, s, 0, 0, 0, s, 0]], np.float64)
        det = s * s
        t = oracle.Tree(tri)
        h, _ = t.closest_hit(_ray((s / 4, s / 4, 1.0), (0, 0, -1)))
        assert (h[0] == 0) == (abs(det) > 1e-7) == hits


def test_tie_goes_to_last_visited(oracle):
    # two coincident triangles: the reference overwrites on equal t, so the
    # one visited last (larger sorted position) wins
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    t = oracle.Tree(np.array([tri, tri], np.float64))
    keys, perm = t.sorted()
    h, _ = t.closest_hit(_ray((0.25, 0.25, 1.0), (0, 0, -1)))
    assert h[0] == perm[-1]


def test_traversal_equals_brute_force(oracle):
    pos = scenes.random_triangles(4000, seed=4, mean_edge=0.03)
    rays = scenes.incoherent_rays(3000, seed=9)
    t = oracle.Tree(pos)
    h1, t1 = t.closest_hit(rays)
    h2, t2 = oracle.closest_hit_brute(pos, rays)
    assert np.array_equal(t1, t2)  # identical FP64 arithmetic per triangle
    assert np.array_equal(h1, h2)  # no exact ties among random triangles
    assert (h1 >= 0).mean() > 0.3


def test_cube_fixture(oracle):
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "cube.obj")
    pos, nrm = scenes.parse_obj(path)
    assert pos.shape == (12, 9) and nrm is not None
    t = oracle.Tree(pos)
    # axis rays from outside hit the faces at distance 1.5 from +-2
    rays = np.concatenate([
        _ray((0, 0, 2), (0, 0, -1)), _ray((0, 0, -2), (0, 0, 1)),
        _ray((2, 0.1, 0.1), (-1, 0, 0)), _ray((0.1, -2, 0.2), (0, 1, 0)),
    ])
    h, tt = t.closest_hit(rays)
    assert np.all(h >= 0)
    assert np.allclose(tt, 1.5)
    # the face hit from +z is one of the two triangles with z = +0.5 everywhere
    assert np.all(pos[h[0]].reshape(3, 3)[:, 2] == 0.5)
