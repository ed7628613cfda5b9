"""ORACLE -- test infrastructure only.

ctypes access to the FP64 CPU restatement in oracle.cpp.  Only tests/, the
smoke check and bench.py's CPU-baseline leg import this module; the product
library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_P = C.c_void_p
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "oracle_split_by3": ([C.c_uint64], C.c_uint64),
        "oracle_enclosing": ([_P, C.c_int64, _P], None),
        "oracle_morton": ([_P, C.c_int64, _P, C.c_int, _P], None),
        "oracle_sort": ([_P, C.c_int64, _P, _P], None),
        "oracle_lbvh_build": ([_P, C.c_int64, _P, C.c_int], _P),
        "oracle_tree_free": ([_P], None),
        "oracle_tree_node_count": ([_P], C.c_int64),
        "oracle_tree_sorted": ([_P, _P, _P], None),
        "oracle_tree_canonical": ([_P, _P, _P], None),
        "oracle_tree_cost": ([_P], C.c_double),
        "oracle_closest_hit": ([_P, _P, C.c_int64, _P, _P, C.c_int], None),
        "oracle_closest_hit_brute": ([_P, C.c_int64, _P, C.c_int64, _P, _P, C.c_int], None),
        "oracle_philox": ([C.c_uint32] * 6 + [_P], None),
        "oracle_phr_build": ([_P, C.c_double, C.c_double, C.c_int, C.c_int64], _P),
        "oracle_phr_build_traced": ([_P, C.c_double, C.c_double, C.c_int, C.c_int64], _P),
        "oracle_phr_trace": ([_P, _P, _P, _P, _P], None),
        "oracle_phr_initial_ranges": ([_P, _P], C.c_int64),
        "oracle_phr_free": ([_P], None),
        "oracle_phr_cost": ([_P], C.c_double),
        "oracle_phr_split_work": ([_P], C.c_int64),
        "oracle_phr_initial_cut": ([_P], C.c_int64),
        "oracle_phr_canonical": ([_P, _P, _P, _P, _P], C.c_int64),
        "oracle_phr_closest_hit": ([_P, _P, C.c_int64, _P, _P, C.c_int], None),
        "oracle_traversal_steps": ([_P, _P, C.c_int64, _P, C.c_int], None),
        "oracle_scene_flatten": ([_P, _P, _P, _P, _P, _P, _P, C.c_int32, _P, _P, _P], C.c_int64),
        "oracle_phr_traversal_steps": ([_P, _P, C.c_int64, _P, C.c_int], None),
        "oracle_heatmap": ([_P, _P, C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_int], None),
        "oracle_phr_heatmap": ([_P, _P, C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_int], None),
        "oracle_render": ([_P, _P, _P, _P, _P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                           C.c_uint64, C.c_int64, C.c_int, _P, _P, _P, _P], None),
        "oracle_render_ex": ([_P, _P, _P, _P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.c_uint64, C.c_int, _P, _P, _P], None),
        "oracle_render_grad": ([_P, _P, _P, _P, _P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                C.c_int32, C.c_uint64, C.c_int64, C.c_int, _P, _P, _P, _P], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def split_by3(a: int) -> int:
    return int(load().oracle_split_by3(a))


def enclosing(pos: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(pos, np.float64)
    out = np.empty(6, np.float64)
    load().oracle_enclosing(_p(p), p.shape[0], _p(out))
    return out


def morton(pos: np.ndarray, qbox: np.ndarray, key_mode: int = 0) -> np.ndarray:
    p = np.ascontiguousarray(pos, np.float64)
    q = np.ascontiguousarray(qbox, np.float64)
    keys = np.empty(p.shape[0], np.uint64)
    load().oracle_morton(_p(p), p.shape[0], _p(q), key_mode, _p(keys))
    return keys


def stable_sort(keys: np.ndarray):
    k = np.ascontiguousarray(keys, np.uint64)
    s = np.empty_like(k)
    perm = np.empty(k.shape[0], np.int32)
    load().oracle_sort(_p(k), k.shape[0], _p(s), _p(perm))
    return s, perm


def _steps(fn, h, rays: np.ndarray, threads: int | None) -> np.ndarray:
    r = np.ascontiguousarray(rays, np.float64).reshape(-1, 8)
    counts = np.empty(r.shape[0], np.int32)
    fn(h, _p(r), r.shape[0], _p(counts), threads or _threads())
    return counts


def _heatmap(fn, h, cam, width: int, height: int, threshold: int, threads: int | None):
    c = np.array([*cam.origin, *cam.lower_left, *cam.horizontal, *cam.vertical], np.float64)
    rgb = np.empty((height, width, 3), np.float32)
    counts = np.empty((height, width), np.int32)
    fn(h, _p(c), width, height, threshold, _p(rgb), _p(counts), threads or _threads())
    return rgb, counts


class Tree:
    """Reference LBVH (top-down split, FP64 boxes) built by the oracle."""

    def __init__(self, pos: np.ndarray, enclosing_box=None, key_mode: int = 0):
        self.lib = load()
        self.pos = np.ascontiguousarray(pos, np.float64)
        enc = None if enclosing_box is None else np.ascontiguousarray(enclosing_box, np.float64)
        self.h = self.lib.oracle_lbvh_build(_p(self.pos), self.pos.shape[0], None if enc is None else _p(enc),
                                            key_mode)
        if not self.h:
            raise ValueError("oracle LBVH needs at least one triangle")
        self.n = self.pos.shape[0]

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_tree_free(self.h)
            self.h = None

    def node_count(self) -> int:
        return int(self.lib.oracle_tree_node_count(self.h))

    def sorted(self):
        keys = np.empty(self.n, np.uint64)
        perm = np.empty(self.n, np.int32)
        self.lib.oracle_tree_sorted(self.h, _p(keys), _p(perm))
        return keys, perm

    def canonical(self):
        n = self.node_count()
        rec = np.empty((n, 4), np.int32)
        box = np.empty((n, 6), np.float64)
        self.lib.oracle_tree_canonical(self.h, _p(rec), _p(box))
        return rec, box

    def cost(self) -> float:
        return float(self.lib.oracle_tree_cost(self.h))

    def closest_hit(self, rays: np.ndarray, threads: int | None = None):
        r = np.ascontiguousarray(rays, np.float32)
        n = r.shape[0]
        hit = np.empty(n, np.int32)
        t = np.empty(n, np.float64)
        self.lib.oracle_closest_hit(self.h, _p(r), n, _p(hit), _p(t), threads or _threads())
        return hit, t

    def traversal_steps(self, rays: np.ndarray, threads: int | None = None) -> np.ndarray:
        """BVH.traversalSteps per FP64 ray (o, tMin, d, tMax)."""
        return _steps(self.lib.oracle_traversal_steps, self.h, rays, threads)

    def heatmap(self, cam, width: int, height: int, threshold: int, threads: int | None = None):
        """HeatMapRenderer.RenderToBuffer: (rgb (H, W, 3) float32, counts (H, W))."""
        return _heatmap(self.lib.oracle_heatmap, self.h, cam, width, height, threshold, threads)


class PhrTree:
    """PhrBuilder.BuildFromAuxilary on an oracle LBVH (deterministic FIFO initial cut)."""

    def __init__(self, aux: Tree, alpha: float = 0.5, delta: float = 6.0, branching: int = 2, max_cut: int = 2500,
                 trace: bool = False):
        self.lib = load()
        self.aux = aux  # keep the LBVH alive
        build = self.lib.oracle_phr_build_traced if trace else self.lib.oracle_phr_build
        self.h = build(aux.h, alpha, delta, branching, max_cut)

    def split_calls(self):
        """The SweepSAH calls of a traced build, in order: a list of dicts with
        'ranges' (m, 2) sorted-position ranges of the members (-1 for a PHR-made
        leaf), 'boxes' (m, 6), 'bounding' (6,), 'split' (bool), 'left' /
        'right' member indices in the order the split cuts hold them, and the
        sizes of the job's cuts at that moment ('cut_sizes') with the index of
        the one split ('chosen')."""
        ni, nd = C.c_int64(), C.c_int64()
        self.lib.oracle_phr_trace(self.h, None, None, C.byref(ni), C.byref(nd))
        ints = np.empty(ni.value, np.int32)
        dbl = np.empty(nd.value, np.float64)
        self.lib.oracle_phr_trace(self.h, _p(ints), _p(dbl), C.byref(ni), C.byref(nd))
        out, i, d = [], 0, 0
        while i < len(ints):
            m, split, nl, nr, ncuts, chosen = (int(v) for v in ints[i:i + 6])
            i += 6
            sizes = ints[i:i + ncuts].copy()
            i += ncuts
            ranges = ints[i:i + 2 * m].reshape(m, 2).copy()
            i += 2 * m
            left = ints[i:i + nl].copy()
            i += nl
            right = ints[i:i + nr].copy()
            i += nr
            bounding = dbl[d:d + 6].copy()
            boxes = dbl[d + 6:d + 6 + 6 * m].reshape(m, 6).copy()
            d += 6 + 6 * m
            out.append({"ranges": ranges, "boxes": boxes, "bounding": bounding, "split": bool(split),
                        "left": left, "right": right, "cut_sizes": sizes, "chosen": chosen})
        return out

    def initial_ranges(self) -> np.ndarray:
        """The initial cut in order: (first, last) sorted-position range per member."""
        n = int(self.lib.oracle_phr_initial_ranges(self.h, None))
        out = np.empty((n, 2), np.int32)
        self.lib.oracle_phr_initial_ranges(self.h, _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_phr_free(self.h)
            self.h = None

    def cost(self) -> float:
        return float(self.lib.oracle_phr_cost(self.h))

    def split_work(self) -> int:
        """BuildWithCost's cost (phr.go:154-255): initial cut + cuts split."""
        return int(self.lib.oracle_phr_split_work(self.h))

    def build_cost(self) -> int:
        return self.split_work()

    def initial_cut(self) -> int:
        return int(self.lib.oracle_phr_initial_cut(self.h))

    def closest_hit(self, rays: np.ndarray, threads: int | None = None):
        r = np.ascontiguousarray(rays, np.float32)
        n = r.shape[0]
        hit = np.empty(n, np.int32)
        t = np.empty(n, np.float64)
        self.lib.oracle_phr_closest_hit(self.h, _p(r), n, _p(hit), _p(t), threads or _threads())
        return hit, t

    def traversal_steps(self, rays: np.ndarray, threads: int | None = None) -> np.ndarray:
        return _steps(self.lib.oracle_phr_traversal_steps, self.h, rays, threads)

    def heatmap(self, cam, width: int, height: int, threshold: int, threads: int | None = None):
        return _heatmap(self.lib.oracle_phr_heatmap, self.h, cam, width, height, threshold, threads)

    def canonical(self):
        """(rec (n,4): kind, prim count or -1, n_children, 0; box (n,6) FP64;
        prims: sorted positions of the leaves in preorder)."""
        npr = C.c_int64()
        n = int(self.lib.oracle_phr_canonical(self.h, None, None, None, C.byref(npr)))
        rec = np.empty((n, 4), np.int32)
        box = np.empty((n, 6), np.float64)
        prims = np.empty(npr.value, np.int32)
        self.lib.oracle_phr_canonical(self.h, _p(rec), _p(box), _p(prims), C.byref(npr))
        return rec, box, prims


def omega(prims: int, pixels: int) -> float:
    """omega (optimizer.go:88-90): (prims / PRIM_THRESH)^2 * pixels / RES_THRESH."""
    r = float(prims) / 2000000
    return (r * r) * (float(pixels) / 200000)


def eval_phr(sah: float, max_sah: float, build: int, max_build: int, om: float) -> float:
    """evalPHR (optimizer.go:92-94)."""
    return float(build) / float(max_build) + (om * om) * (sah / max_sah)


def bayes_objective(tree: Tree, alpha: float, delta: float, alpha_range, delta_range, width: int, height: int,
                    branching: int = 2, max_cut: int = 2500) -> tuple:
    """The objective bayesianOptimizer.OptimizedPHRparams hands to its search
    (optimizer.go:132-151) at one (alpha, delta): Cost() of the tree at the range
    minimum and the build cost at the range maximum normalise evalPHR; a zero
    parameter scores the normalisers.  Returns (objective, Cost, build cost)."""
    max_sah = PhrTree(tree, alpha_range[0], delta_range[0], branching, max_cut).cost()
    max_build = PhrTree(tree, alpha_range[1], delta_range[1], branching, max_cut).build_cost()
    sah, build = max_sah, max_build
    if alpha != 0 and delta != 0:
        p = PhrTree(tree, alpha, delta, branching, max_cut)
        sah, build = p.cost(), p.build_cost()
    return eval_phr(sah, max_sah, build, max_build, omega(tree.n, width * height)), sah, build


def grid_optimize(tree: Tree, alphas, deltas, width: int, height: int, branching: int = 2,
                  max_cut: int = 2500):
    """gridOptimizer.OptimizedPHRparams (optimizer.go:44-94): BuildWithCost
    (phr.go:154-255) for every (alpha, delta), alphas outer; each evaluation
    scores buildCost / max buildCost + omega^2 * SAH / max SAH with omega =
    (prims / PRIM_THRESH)^2 * pixels / RES_THRESH (optimizer.go:10-13, 87-94);
    the first minimum wins.  Returns (alpha, delta, [(sah, build_cost)...])."""
    evals = []
    for a in alphas:
        for d in deltas:
            p = PhrTree(tree, a, d, branching, max_cut)
            evals.append((p.cost(), p.build_cost(), float(a), float(d)))
    max_build = max(e[1] for e in evals)
    max_sah = max(e[0] for e in evals)
    om = omega(tree.n, width * height)
    best, best_cost = None, np.finfo(np.float64).max
    for sah, build, a, d in evals:
        c = eval_phr(sah, max_sah, build, max_build, om)
        if c < best_cost:
            best_cost, best = c, (a, d)
    return best[0], best[1], [(e[0], e[1]) for e in evals]


def scene_flatten(meshes, mesh_materials, parent, node_mesh, matrices, normals=None):
    """Scene.collectTracables (scene.go:95-151, primitive.go:152-170) in FP64:
    (positions (n, 9), normals (n, 9) or None, material per triangle)."""
    lib = load()
    parts = [np.ascontiguousarray(m, np.float64).reshape(-1, 9) for m in meshes]
    pos = np.ascontiguousarray(np.concatenate(parts))
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in parts])
    nrm = None
    if normals is not None:
        nrm = np.ascontiguousarray(np.concatenate([np.asarray(n, np.float64).reshape(-1, 9) for n in normals]))
    mm = np.ascontiguousarray(mesh_materials, np.int32)
    par = np.ascontiguousarray(parent, np.int32)
    nm = np.ascontiguousarray(node_mesh, np.int32)
    mx = np.ascontiguousarray(matrices, np.float64).reshape(-1, 16)
    args = [_p(pos), None if nrm is None else _p(nrm), _p(off), _p(mm), _p(par), _p(nm), _p(mx), len(par)]
    n = int(lib.oracle_scene_flatten(*args, None, None, None))
    out_pos = np.empty((n, 9), np.float64)
    out_nrm = np.empty((n, 9), np.float64) if nrm is not None else None
    out_mat = np.empty(n, np.int32)
    lib.oracle_scene_flatten(*args, _p(out_pos), None if out_nrm is None else _p(out_nrm), _p(out_mat))
    return out_pos, out_nrm, out_mat


def philox(counter, key) -> list[int]:
    out = np.empty(4, np.uint32)
    load().oracle_philox(*[int(c) for c in counter], *[int(k) for k in key], _p(out))
    return [int(v) for v in out]


def _opt_f32(a, n: int):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, np.float32).reshape(-1)
    if arr.shape[0] != 3 * n:
        raise ValueError("per-triangle parameters need 3 floats per triangle")
    return arr, _p(arr)


def render(tree: Tree, mat_idx, materials, cam, width: int, height: int, spp: int, max_depth: int,
           miss: int, seed: int, sample_offset: int = 0, normals=None, path_values: bool = False,
           threads: int | None = None, albedo=None, emission=None):
    """FP64 restatement of RenderToBuffer.  cam has origin, lower_left,
    horizontal, vertical (scenes.Camera).  albedo / emission (optional, (n, 3))
    replace the material colour of Diffuse / Light triangles.  Returns
    (sum, sumsq, last[, paths]), each (height, width, 3) float64 (paths:
    (spp, height, width, 3))."""
    mi = np.ascontiguousarray(mat_idx, np.int32)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float64)
    if np.asarray(materials).reshape(len(materials), -1).shape[1] == 8:
        # (kind, r, g, b, param, 0, 0, 0): Reflective / Refractive too
        m8 = np.ascontiguousarray(materials, np.float32).reshape(-1, 8)
        c = np.ascontiguousarray(np.concatenate([cam.origin, cam.lower_left, cam.horizontal, cam.vertical]),
                                 np.float64)
        shape = (height, width, 3)
        s, q, last = (np.empty(shape, np.float64) for _ in range(3))
        load().oracle_render_ex(tree.h, None if nrm is None else _p(nrm), _p(mi), _p(m8), m8.shape[0], _p(c), width,
                                height, spp, max_depth, miss, seed, threads or _threads(), _p(s), _p(q), _p(last))
        return s, q, last
    mats = np.ascontiguousarray(materials, np.float32).reshape(-1, 4)
    alb, palb = _opt_f32(albedo, len(mi))
    emi, pemi = _opt_f32(emission, len(mi))
    c = np.ascontiguousarray(np.concatenate([cam.origin, cam.lower_left, cam.horizontal, cam.vertical]), np.float64)
    shape = (height, width, 3)
    s = np.empty(shape, np.float64)
    q = np.empty(shape, np.float64)
    last = np.empty(shape, np.float64)
    paths = np.empty((spp, *shape), np.float64) if path_values else None
    load().oracle_render(tree.h, None if nrm is None else _p(nrm), _p(mi), _p(mats), palb, pemi, _p(c), width,
                         height, spp, max_depth, miss, seed, sample_offset, threads or _threads(), _p(s), _p(q),
                         _p(last), None if paths is None else _p(paths))
    return (s, q, last, paths) if path_values else (s, q, last)


def render_grad(tree: Tree, mat_idx, materials, cam, width: int, height: int, spp: int, max_depth: int,
                miss: int, seed: int, g_img, sample_offset: int = 0, normals=None, threads: int | None = None,
                albedo=None, emission=None):
    """FP64 gradient of a loss on the mean image with respect to per-triangle
    albedo and emission, given g_img = dLoss/d(mean image) (height, width, 3).
    Returns (mean image, g_albedo (n, 3), g_emission (n, 3))."""
    mi = np.ascontiguousarray(mat_idx, np.int32)
    n = len(mi)
    mats = np.ascontiguousarray(materials, np.float32).reshape(-1, 4)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float64)
    alb, palb = _opt_f32(albedo, n)
    emi, pemi = _opt_f32(emission, n)
    c = np.ascontiguousarray(np.concatenate([cam.origin, cam.lower_left, cam.horizontal, cam.vertical]), np.float64)
    g = np.ascontiguousarray(g_img, np.float64).reshape(height, width, 3)
    mean = np.empty((height, width, 3), np.float64)
    ga = np.empty((n, 3), np.float64)
    ge = np.empty((n, 3), np.float64)
    load().oracle_render_grad(tree.h, None if nrm is None else _p(nrm), _p(mi), _p(mats), palb, pemi, _p(c), width,
                              height, spp, max_depth, miss, seed, sample_offset, threads or _threads(), _p(g),
                              _p(mean), _p(ga), _p(ge))
    return mean, ga, ge


def closest_hit_brute(pos: np.ndarray, rays: np.ndarray, threads: int | None = None):
    p = np.ascontiguousarray(pos, np.float64)
    r = np.ascontiguousarray(rays, np.float32)
    n = r.shape[0]
    hit = np.empty(n, np.int32)
    t = np.empty(n, np.float64)
    load().oracle_closest_hit_brute(_p(p), p.shape[0], _p(r), n, _p(hit), _p(t), threads or _threads())
    return hit, t
