"""Thin object layer over the C-ABI: contexts, device meshes and BVH handles.

All arrays cross as numpy buffers (host) or raw device addresses (ints).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BvhInfo, check, load


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Context:
    """A HIP device context (one stream).  ``trace_ctx_create``."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.trace_ctx_create(device, C.byref(h))
        if rc != _lib.TRACE_OK:
            raise _lib.TraceError(rc, f"trace_ctx_create(device={device}) failed")
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.trace_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int) -> None:
        check(rc, self.h)

    def set_stream(self, stream_ptr: int | None) -> None:
        self.check(self.lib.trace_ctx_set_stream(self.h, C.c_void_p(stream_ptr or 0)))

    def synchronize(self) -> None:
        self.check(self.lib.trace_ctx_synchronize(self.h))


class Mesh:
    """Device-resident triangles (9 doubles each), ``trace_mesh_create``."""

    def __init__(self, ctx: Context, pos: np.ndarray | None = None, *, device_ptr: int | None = None,
                 ntri: int | None = None):
        self.ctx = ctx
        h = C.c_void_p()
        if device_ptr is not None:
            self.ntri = int(ntri)
            ctx.check(ctx.lib.trace_mesh_create_device(ctx.h, C.c_void_p(device_ptr), self.ntri, C.byref(h)))
        else:
            p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 9)
            self.ntri = p.shape[0]
            ctx.check(ctx.lib.trace_mesh_create(ctx.h, _ptr(p), self.ntri, C.byref(h)))
        self.h = h

    @classmethod
    def segmented(cls, ctx: Context, meshes) -> "Mesh":
        """Many meshes in one buffer (trace_mesh_create_segmented); meshes is a
        list of (n_i, 9) arrays.  Its LBVH is a forest with one tree per mesh."""
        parts = [np.ascontiguousarray(m, np.float64).reshape(-1, 9) for m in meshes]
        pos = np.ascontiguousarray(np.concatenate(parts))
        off = np.zeros(len(parts) + 1, np.int64)
        off[1:] = np.cumsum([len(p) for p in parts])
        self = cls.__new__(cls)
        self.ctx = ctx
        self.ntri = pos.shape[0]
        self.offsets = off
        h = C.c_void_p()
        ctx.check(ctx.lib.trace_mesh_create_segmented(ctx.h, _ptr(pos), self.ntri, _ptr(off), len(parts),
                                                      C.byref(h)))
        self.h = h
        return self

    def export(self) -> np.ndarray:
        out = np.empty((self.ntri, 9), np.float64)
        self.ctx.check(self.ctx.lib.trace_mesh_export(self.ctx.h, self.h, _ptr(out)))
        return out

    def animate(self, rest: "Mesh", amplitude: float, period: float, seed: int, d_frame: int) -> None:
        """trace_mesh_animate (asynchronous; reads and increments *d_frame)."""
        self.ctx.check(self.ctx.lib.trace_mesh_animate(self.ctx.h, self.h, rest.h, amplitude, period, seed,
                                                       C.c_void_p(d_frame)))

    def bind_params(self, d_albedo: int | None, d_emission: int | None) -> None:
        """trace_mesh_bind_params: per-triangle albedo / emission (device
        pointers to 3 floats per triangle, kept alive by the caller); None
        unbinds."""
        self.ctx.check(self.ctx.lib.trace_mesh_bind_params(self.ctx.h, self.h, C.c_void_p(d_albedo or 0),
                                                           C.c_void_p(d_emission or 0)))

    def set_shading(self, mat_idx, materials, normals=None) -> None:
        """Per-triangle material index, material table (kind, r, g, b) and
        optional vertex normals (n, 9); None = geometric normals."""
        mi = np.ascontiguousarray(mat_idx, np.int32)
        mats = np.ascontiguousarray(materials, np.float32).reshape(-1, 4)
        if mi.shape[0] != self.ntri:
            raise ValueError("one material index per triangle")
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float64).reshape(-1, 9)
        self.ctx.check(self.ctx.lib.trace_mesh_set_shading(self.ctx.h, self.h, None if nrm is None else _ptr(nrm),
                                                           _ptr(mi), _ptr(mats), mats.shape[0]))

    def set_shading_ex(self, mat_idx, materials8, normals=None) -> None:
        """All four material kinds: materials8 rows are (kind, r, g, b, param,
        0, 0, 0) with param = Reflective.Diffusion or Refractive.Ratio."""
        mi = np.ascontiguousarray(mat_idx, np.int32)
        mats = np.ascontiguousarray(materials8, np.float32).reshape(-1, 8)
        if mi.shape[0] != self.ntri:
            raise ValueError("one material index per triangle")
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float64).reshape(-1, 9)
        self.ctx.check(self.ctx.lib.trace_mesh_set_shading_ex(self.ctx.h, self.h, None if nrm is None else _ptr(nrm),
                                                              _ptr(mi), _ptr(mats), mats.shape[0]))

    def export_shading(self):
        """(normals (n, 9) float32, material index (n,)) as the renderer holds them."""
        nrm = np.empty((self.ntri, 9), np.float32)
        mi = np.empty(self.ntri, np.int32)
        self.ctx.check(self.ctx.lib.trace_mesh_export_shading(self.ctx.h, self.h, _ptr(nrm), _ptr(mi)))
        return nrm, mi

    def update(self, pos: np.ndarray) -> None:
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 9)
        if p.shape[0] != self.ntri:
            raise ValueError("update must keep the triangle count")
        self.ctx.check(self.ctx.lib.trace_mesh_update(self.ctx.h, self.h, _ptr(p)))

    def close(self) -> None:
        if self.h:
            self.ctx.lib.trace_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Meshes of a scene graph on the device (trace_scene_create); compile()
    flattens a node tree into a Mesh with shading (trace_scene_compile)."""

    def __init__(self, ctx: Context, meshes, mesh_materials, materials, normals=None):
        self.ctx = ctx
        parts = [np.ascontiguousarray(m, np.float64).reshape(-1, 9) for m in meshes]
        pos = np.ascontiguousarray(np.concatenate(parts))
        off = np.zeros(len(parts) + 1, np.int64)
        off[1:] = np.cumsum([len(p) for p in parts])
        nrm = None
        if normals is not None:
            nrm = np.ascontiguousarray(np.concatenate([np.asarray(n, np.float64).reshape(-1, 9) for n in normals]))
        mm = np.ascontiguousarray(mesh_materials, np.int32)
        mats = np.ascontiguousarray(materials, np.float32).reshape(-1, 4)
        self.sizes = [len(p) for p in parts]
        h = C.c_void_p()
        ctx.check(ctx.lib.trace_scene_create(ctx.h, _ptr(pos), None if nrm is None else _ptr(nrm), _ptr(off),
                                             len(parts), _ptr(mm), _ptr(mats), mats.shape[0], C.byref(h)))
        self.h = h

    def compile(self, parent, node_mesh, matrices, mesh: Mesh | None = None) -> Mesh:
        """Scene.collectTracables on the device; matrices (n, 16) row-major.
        Refills `mesh` in place when given (same triangle count)."""
        par = np.ascontiguousarray(parent, np.int32)
        nm = np.ascontiguousarray(node_mesh, np.int32)
        mx = np.ascontiguousarray(matrices, np.float64).reshape(-1, 16)
        h = C.c_void_p(mesh.h.value if mesh is not None else None)
        self.ctx.check(self.ctx.lib.trace_scene_compile(self.ctx.h, self.h, _ptr(par), _ptr(nm), _ptr(mx), len(par),
                                                        C.byref(h)))
        if mesh is not None:
            return mesh
        m = Mesh.__new__(Mesh)
        m.ctx = self.ctx
        m.h = h
        m.ntri = int(sum(self.sizes[k] for k in nm if k >= 0))  # every mesh occurrence is emitted
        return m

    def close(self) -> None:
        if self.h:
            self.ctx.lib.trace_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Bvh:
    """A device BVH handle."""

    def __init__(self, ctx: Context, handle: C.c_void_p, n_prims: int):
        self.ctx = ctx
        self.h = handle
        self.n_prims = n_prims

    @classmethod
    def lbvh(cls, ctx: Context, mesh: Mesh, enclosing=None, key_mode: int = _lib.KEY_REF63) -> "Bvh":
        h = C.c_void_p()
        enc = None if enclosing is None else np.ascontiguousarray(enclosing, dtype=np.float64)
        ctx.check(ctx.lib.trace_lbvh_build(ctx.h, mesh.h, None if enc is None else _ptr(enc), key_mode,
                                           C.byref(h)))
        b = cls(ctx, h, mesh.ntri)
        b._enc = enc
        return b

    @classmethod
    def phr(cls, ctx: Context, aux: "Bvh", alpha: float = 0.5, delta: float = 6.0, branching: int = 2,
            max_cut: int = 2500) -> "Bvh":
        """PhrBuilder.BuildFromAuxilary on an LBVH (trace_phr_build)."""
        h = C.c_void_p()
        ctx.check(ctx.lib.trace_phr_build(ctx.h, aux.h, alpha, delta, branching, max_cut, C.byref(h)))
        b = cls(ctx, h, aux.n_prims)
        b._aux = aux
        return b

    def phr_rebuild(self, aux: "Bvh", alpha: float = 0.5, delta: float = 6.0, branching: int = 2,
                    max_cut: int = 2500) -> None:
        """Rebuilds this PHR tree from aux in place (trace_phr_rebuild)."""
        self.ctx.check(self.ctx.lib.trace_phr_rebuild(self.ctx.h, aux.h, alpha, delta, branching, max_cut, self.h))
        self.n_prims = aux.n_prims
        self._aux = aux

    @classmethod
    def phr_with_cost(cls, ctx: Context, aux: "Bvh", alpha: float = 0.5, delta: float = 6.0, branching: int = 2,
                      max_cut: int = 2500):
        """PhrBuilder.BuildWithCost: (tree, build cost) (trace_phr_build_with_cost)."""
        h = C.c_void_p()
        cost = C.c_int64()
        ctx.check(ctx.lib.trace_phr_build_with_cost(ctx.h, aux.h, alpha, delta, branching, max_cut, C.byref(h),
                                                    C.byref(cost)))
        b = cls(ctx, h, aux.n_prims)
        b._aux = aux
        return b, int(cost.value)

    @classmethod
    def instances(cls, ctx: Context, blas, inst_blas, xforms, tri_base) -> "Bvh":
        """trace_instances_build: instance i places blas[inst_blas[i]] with the affine
        object-to-world transform xforms[i] (3 x 4 rows (A | t)); its triangles are
        numbered from tri_base[i] in the flattened order."""
        arr = (C.c_void_p * len(blas))(*[b.h.value for b in blas])
        ib = np.ascontiguousarray(inst_blas, np.int32)
        xf = np.ascontiguousarray(xforms, np.float64).reshape(-1, 12)
        tb = np.ascontiguousarray(tri_base, np.int64)
        h = C.c_void_p()
        ctx.check(ctx.lib.trace_instances_build(ctx.h, arr, len(blas), _ptr(ib), _ptr(xf), _ptr(tb), len(ib),
                                                C.byref(h)))
        b = cls(ctx, h, int(max(int(tb[i]) + blas[int(ib[i])].n_prims for i in range(len(ib)))))
        b._blas = list(blas)  # the bottom-level trees must outlive the instanced one
        return b

    def device_bytes(self) -> int:
        """trace_bvh_device_bytes."""
        v = C.c_int64()
        self.ctx.check(self.ctx.lib.trace_bvh_device_bytes(self.ctx.h, self.h, C.byref(v)))
        return int(v.value)

    def segment(self, seg: int) -> "Bvh":
        """View of tree `seg` of a forest (trace_bvh_segment)."""
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.trace_bvh_segment(self.ctx.h, self.h, seg, C.byref(h)))
        v = Bvh(self.ctx, h, 0)
        v._forest = self  # keep the forest alive while the view exists
        v.n_prims = int(v.info().n_prims)
        return v

    def phr_source(self) -> np.ndarray:
        src = np.empty(self.n_prims, np.int32)
        self.ctx.check(self.ctx.lib.trace_phr_export_source(self.ctx.h, self.h, _ptr(src)))
        return src

    def rebuild(self, mesh: Mesh, enclosing=None, key_mode: int = _lib.KEY_REF63) -> None:
        """Asynchronous rebuild into this handle (no device allocation once sized)."""
        enc = None if enclosing is None else np.ascontiguousarray(enclosing, dtype=np.float64)
        self._enc = enc
        self.ctx.check(self.ctx.lib.trace_lbvh_rebuild(self.ctx.h, mesh.h, None if enc is None else _ptr(enc),
                                                       key_mode, self.h))

    def rebuild_noalloc(self, mesh: Mesh, enclosing=None, key_mode: int = _lib.KEY_REF63,
                        ctx: Context | None = None) -> None:
        """trace_lbvh_rebuild_noalloc (asynchronous): the rebuild with the
        library's fixed scratch array, on `ctx` (default: the tree's context)."""
        c = ctx or self.ctx
        enc = None if enclosing is None else np.ascontiguousarray(enclosing, dtype=np.float64)
        self._enc = enc
        c.check(c.lib.trace_lbvh_rebuild_noalloc(c.h, mesh.h, None if enc is None else _ptr(enc), key_mode, self.h))

    def close(self) -> None:
        if self.h:
            self.ctx.lib.trace_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> BvhInfo:
        out = BvhInfo()
        self.ctx.check(self.ctx.lib.trace_bvh_get_info(self.ctx.h, self.h, C.byref(out)))
        return out

    def sorted(self):
        keys = np.empty(self.n_prims, np.uint64)
        perm = np.empty(self.n_prims, np.int32)
        self.ctx.check(self.ctx.lib.trace_bvh_export_sorted(self.ctx.h, self.h, _ptr(keys), _ptr(perm)))
        return keys, perm

    def perm(self) -> np.ndarray:
        """Triangle index of each position in the tree's triangle order."""
        perm = np.empty(self.n_prims, np.int32)
        self.ctx.check(self.ctx.lib.trace_bvh_export_sorted(self.ctx.h, self.h, None, _ptr(perm)))
        return perm

    def canonical(self):
        n = C.c_int64()
        self.ctx.check(self.ctx.lib.trace_bvh_export_canonical(self.ctx.h, self.h, 0, None, None, C.byref(n)))
        rec = np.empty((n.value, 4), np.int32)
        box = np.empty((n.value, 6), np.float32)
        self.ctx.check(self.ctx.lib.trace_bvh_export_canonical(self.ctx.h, self.h, n.value, _ptr(rec), _ptr(box),
                                                               C.byref(n)))
        return rec, box

    def cost(self) -> float:
        c = C.c_double()
        self.ctx.check(self.ctx.lib.trace_bvh_cost(self.ctx.h, self.h, C.byref(c)))
        return c.value

    def size(self) -> int:
        s = C.c_int64()
        self.ctx.check(self.ctx.lib.trace_bvh_size(self.ctx.h, self.h, C.byref(s)))
        return s.value

    def closest_hit(self, rays: np.ndarray):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = r.shape[0]
        hit = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        self.ctx.check(self.ctx.lib.trace_closest_hit(self.ctx.h, self.h, _ptr(r), n, _ptr(hit), _ptr(t)))
        return hit, t

    def closest_hit_device(self, d_rays: int, n: int, d_hit: int, d_t: int) -> None:
        self.ctx.check(self.ctx.lib.trace_closest_hit_device(self.ctx.h, self.h, C.c_void_p(d_rays), n,
                                                             C.c_void_p(d_hit), C.c_void_p(d_t)))

    def traversal_steps(self, rays: np.ndarray) -> np.ndarray:
        """BVH.traversalSteps per FP64 ray (origin, tMin, direction, tMax)."""
        r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 8)
        out = np.empty(r.shape[0], np.int32)
        self.ctx.check(self.ctx.lib.trace_traversal_steps(self.ctx.h, self.h, _ptr(r), r.shape[0], _ptr(out)))
        return out

    def heatmap(self, cam, width: int, height: int, threshold: int):
        """HeatMapRenderer.RenderToBuffer: (rgb (H, W, 3) float32, counts (H, W) int32)."""
        rgb = np.empty((height, width, 3), np.float32)
        counts = np.empty((height, width), np.int32)
        c = camera_struct(cam)
        self.ctx.check(self.ctx.lib.trace_render_heatmap(self.ctx.h, self.h, C.byref(c), width, height, threshold,
                                                         _ptr(rgb), _ptr(counts)))
        return rgb, counts

    def counts(self, rays: np.ndarray):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = r.shape[0]
        nodes = np.empty(n, np.int32)
        tris = np.empty(n, np.int32)
        self.ctx.check(self.ctx.lib.trace_closest_hit_counts(self.ctx.h, self.h, _ptr(r), n, _ptr(nodes),
                                                             _ptr(tris)))
        return nodes, tris


def camera_struct(cam) -> "_lib.Camera":
    c = _lib.Camera()
    for name in ("origin", "lower_left", "horizontal", "vertical"):
        getattr(c, name)[:] = [float(v) for v in getattr(cam, name)]
    return c


def render_params(width: int, height: int, spp: int, max_depth: int, miss: int, seed: int,
                  sample_offset: int = 0) -> "_lib.RenderParams":
    p = _lib.RenderParams()
    p.width, p.height, p.spp, p.max_depth, p.miss = width, height, spp, max_depth, miss
    p.seed, p.sample_offset = seed, sample_offset
    return p


def render(ctx: Context, mesh: Mesh, bvh: Bvh, cam, width: int, height: int, spp: int, max_depth: int,
           miss: int, seed: int, sample_offset: int = 0):
    """trace_render: returns (sum, sumsq, last), each (height, width, 3) float32."""
    shape = (height, width, 3)
    s = np.empty(shape, np.float32)
    q = np.empty(shape, np.float32)
    last = np.empty(shape, np.float32)
    c = camera_struct(cam)
    p = render_params(width, height, spp, max_depth, miss, seed, sample_offset)
    ctx.check(ctx.lib.trace_render(ctx.h, mesh.h, bvh.h, C.byref(c), C.byref(p), _ptr(s), _ptr(q), _ptr(last)))
    return s, q, last


def render_device(ctx: Context, mesh: Mesh, bvh: Bvh, cam, width: int, height: int, spp: int, max_depth: int,
                  miss: int, seed: int, d_sum: int, d_sumsq: int, d_last: int | None = None,
                  sample_offset: int = 0) -> None:
    """trace_render_device (asynchronous on the context stream)."""
    c = camera_struct(cam)
    p = render_params(width, height, spp, max_depth, miss, seed, sample_offset)
    ctx.check(ctx.lib.trace_render_device(ctx.h, mesh.h, bvh.h, C.byref(c), C.byref(p), C.c_void_p(d_sum),
                                          C.c_void_p(d_sumsq), C.c_void_p(d_last or 0)))


def render_backward(ctx: Context, mesh: Mesh, bvh: Bvh, cam, width: int, height: int, spp: int, max_depth: int,
                    miss: int, seed: int, d_grad_image: int, d_grad_albedo: int, d_grad_emission: int,
                    sample_offset: int = 0) -> None:
    """trace_render_backward (asynchronous): dLoss/d(per-triangle albedo,
    emission) from dLoss/d(mean image)."""
    c = camera_struct(cam)
    p = render_params(width, height, spp, max_depth, miss, seed, sample_offset)
    ctx.check(ctx.lib.trace_render_backward(ctx.h, mesh.h, bvh.h, C.byref(c), C.byref(p), C.c_void_p(d_grad_image),
                                            C.c_void_p(d_grad_albedo), C.c_void_p(d_grad_emission)))


def render_keep(ctx: Context, mesh: Mesh, bvh: Bvh, cam, width: int, height: int, spp: int, max_depth: int,
                miss: int, seed: int, d_sum: int, d_sumsq: int, d_keep: int, d_last: int | None = None,
                sample_offset: int = 0) -> None:
    """trace_render_keep (asynchronous): render_device that also keeps the path
    summaries (d_keep: width * height * spp * 4 floats) for render_backward_kept."""
    c = camera_struct(cam)
    p = render_params(width, height, spp, max_depth, miss, seed, sample_offset)
    ctx.check(ctx.lib.trace_render_keep(ctx.h, mesh.h, bvh.h, C.byref(c), C.byref(p), C.c_void_p(d_sum),
                                        C.c_void_p(d_sumsq), C.c_void_p(d_last or 0), C.c_void_p(d_keep)))


def render_backward_kept(ctx: Context, mesh: Mesh, bvh: Bvh, cam, width: int, height: int, spp: int,
                         max_depth: int, miss: int, seed: int, d_keep: int, d_grad_image: int, d_grad_albedo: int,
                         d_grad_emission: int, sample_offset: int = 0) -> None:
    """trace_render_backward_kept (asynchronous): render_backward from the
    summaries of a render_keep call with the same arguments (replay only)."""
    c = camera_struct(cam)
    p = render_params(width, height, spp, max_depth, miss, seed, sample_offset)
    ctx.check(ctx.lib.trace_render_backward_kept(ctx.h, mesh.h, bvh.h, C.byref(c), C.byref(p), C.c_void_p(d_keep),
                                                 C.c_void_p(d_grad_image), C.c_void_p(d_grad_albedo),
                                                 C.c_void_p(d_grad_emission)))


def render_counts(ctx: Context, mesh: Mesh, bvh: Bvh, cam, width: int, height: int, spp: int, max_depth: int,
                  miss: int, seed: int):
    """trace_render_counts: (segments, wide-node visits, primitive tests) of a render."""
    c = camera_struct(cam)
    p = render_params(width, height, spp, max_depth, miss, seed)
    out = np.zeros(3, np.int64)
    ctx.check(ctx.lib.trace_render_counts(ctx.h, mesh.h, bvh.h, C.byref(c), C.byref(p), _ptr(out)))
    return int(out[0]), int(out[1]), int(out[2])


def render_segments(ctx: Context) -> int:
    """trace_render_segments: rays traced by the context's last render or
    backward call (waits for the context stream)."""
    out = C.c_int64()
    ctx.check(ctx.lib.trace_render_segments(ctx.h, C.byref(out)))
    return out.value


def phr_optimize_grid(ctx: Context, aux: Bvh, alphas=(0.4, 0.45, 0.5, 0.55), deltas=(6.0, 7.0, 8.0, 9.0),
                      width: int = 1920, height: int = 1080, branching: int = 2, max_cut: int = 2500):
    """gridOptimizer.OptimizedPHRparams (trace_phr_optimize_grid): (alpha, delta,
    SAH costs, build costs), evaluations alphas-major."""
    a = np.ascontiguousarray(alphas, np.float64)
    d = np.ascontiguousarray(deltas, np.float64)
    sah = np.empty(len(a) * len(d), np.float64)
    build = np.empty(len(a) * len(d), np.int64)
    ao, do = C.c_double(), C.c_double()
    ctx.check(ctx.lib.trace_phr_optimize_grid(ctx.h, aux.h, _ptr(a), len(a), _ptr(d), len(d), branching, max_cut,
                                              width, height, _ptr(sah), _ptr(build), C.byref(ao), C.byref(do)))
    return ao.value, do.value, sah, build


def phr_optimize_bayes(ctx: Context, aux: Bvh, alpha_range=(0.4, 0.55), delta_range=(6.0, 9.0),
                       width: int = 1920, height: int = 1080, branching: int = 2, max_cut: int = 2500,
                       rounds: int = 20, random_rounds: int = 5, seed: int = 1):
    """bayesianOptimizer.OptimizedPHRparams (trace_phr_optimize_bayes): (alpha, delta,
    trace) with trace rows (alpha, delta, Cost, build cost, objective): the two
    normaliser builds, then the evaluations in order."""
    ar = np.ascontiguousarray(alpha_range, np.float64)
    dr = np.ascontiguousarray(delta_range, np.float64)
    if ar.shape != (2,) or dr.shape != (2,):
        raise ValueError("alpha_range and delta_range are (min, max) pairs")
    trace = np.empty((rounds + 2, 5), np.float64)
    ao, do = C.c_double(), C.c_double()
    ctx.check(ctx.lib.trace_phr_optimize_bayes(ctx.h, aux.h, _ptr(ar), _ptr(dr), branching, max_cut, width, height,
                                               rounds, random_rounds, seed, _ptr(trace), C.byref(ao), C.byref(do)))
    return ao.value, do.value, trace


def bayes_minimize(objective, lo, hi, rounds: int = 20, random_rounds: int = 5, seed: int = 1):
    """trace_bayes_minimize (host only): minimises objective(x) -> float over the box
    [lo, hi].  Returns (x_best, xs (rounds, dims), ys (rounds,)).  An exception
    raised by the objective ends the search and is raised again here."""
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    if lo.ndim != 1 or lo.shape != hi.shape:
        raise ValueError("lo and hi must be vectors of one length")
    dims = len(lo)
    xs = np.empty((rounds, dims), np.float64)
    ys = np.empty(rounds, np.float64)
    best = np.empty(dims, np.float64)
    failure = []

    def call(_user, x, y):
        try:
            y[0] = float(objective(np.array([x[k] for k in range(dims)])))
            return _lib.TRACE_OK
        except BaseException as e:  # carried across the C frame
            failure.append(e)
            return _lib.TRACE_E_INVALID

    fn = _lib.OBJECTIVE_FN(call)
    rc = _lib.load().trace_bayes_minimize(_ptr(lo), _ptr(hi), dims, rounds, random_rounds, seed, fn, None,
                                          _ptr(xs), _ptr(ys), _ptr(best))
    if failure:
        raise failure[0]
    if rc != _lib.TRACE_OK:
        raise _lib.TraceError(rc, "trace_bayes_minimize: bad arguments or a NaN objective")
    return best, xs, ys


def adam_step(ctx: Context, d_param: int, d_grad: int, d_m: int, d_v: int, n: int, lr: float, step: int,
              beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, lo: float = 0.0,
              hi: float = 1.0) -> None:
    """trace_adam_step (asynchronous, in place on device buffers of n floats)."""
    ctx.check(ctx.lib.trace_adam_step(ctx.h, C.c_void_p(d_param), C.c_void_p(d_grad), C.c_void_p(d_m),
                                      C.c_void_p(d_v), n, lr, beta1, beta2, eps, step, lo, hi))


def render_batched(ctx: Context, mesh: Mesh, forest: Bvh, seg_ids, cams, width: int, height: int, spp: int,
                   max_depth: int, miss: int, seed: int, d_sum: int, d_sumsq: int, d_last: int | None = None) -> None:
    """trace_render_batched: scene c traces segment seg_ids[c] with cams[c] and seed + c."""
    segs = np.ascontiguousarray(seg_ids, np.int32)
    carr = (_lib.Camera * len(cams))(*[camera_struct(c) for c in cams])
    p = render_params(width, height, spp, max_depth, miss, seed)
    ctx.check(ctx.lib.trace_render_batched(ctx.h, mesh.h, forest.h, len(cams), _ptr(segs), carr, C.byref(p),
                                           C.c_void_p(d_sum), C.c_void_p(d_sumsq), C.c_void_p(d_last or 0)))


class Graph:
    """A HIP graph captured from the context stream (trace_graph_*)."""

    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.h = h

    @classmethod
    def capture(cls, ctx: Context, fn) -> "Graph":
        ctx.check(ctx.lib.trace_graph_begin(ctx.h))
        h = C.c_void_p()
        try:
            fn()
        except BaseException:
            # close the capture; a failed recording is dropped
            if ctx.lib.trace_graph_end(ctx.h, C.byref(h)) == _lib.TRACE_OK and h:
                ctx.lib.trace_graph_destroy(h)
            raise
        ctx.check(ctx.lib.trace_graph_end(ctx.h, C.byref(h)))
        return cls(ctx, h)

    def launch(self) -> None:
        self.ctx.check(self.ctx.lib.trace_graph_launch(self.ctx.h, self.h))

    def close(self) -> None:
        if self.h:
            self.ctx.lib.trace_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_tune(ctx: Context) -> "_lib.Tune":
    t = _lib.Tune()
    ctx.check(ctx.lib.trace_get_tune(ctx.h, C.byref(t)))
    return t


def set_tune(ctx: Context, trav_threads: int = 64, render_threads: int = 64, render_batch: int = 1 << 27,
             trav_persistent: bool = True, render_persistent: bool = False, sort_bits: int = 8,
             sort_keys_per_thread: int = 16, trav_order: bool = True, render_order: int = 1) -> None:
    t = _lib.Tune()
    t.trav_threads, t.render_threads, t.render_batch = trav_threads, render_threads, render_batch
    t.trav_persistent = 1 if trav_persistent else 0
    t.render_persistent = 1 if render_persistent else 0
    t.sort_bits, t.sort_keys_per_thread = sort_bits, sort_keys_per_thread
    t.trav_order = 1 if trav_order else 0
    t.render_order = render_order
    ctx.check(ctx.lib.trace_set_tune(ctx.h, C.byref(t)))


# the key-sort configurations trace_autotune_build times, in its order
SORT_CONFIGS = ((8, 16), (8, 8), (8, 4), (10, 8), (10, 4), (11, 4))


def autotune_build(ctx: Context, mesh: Mesh, key_mode: int = 0) -> np.ndarray:
    """Times the LBVH build for each key-sort configuration (SORT_CONFIGS); keeps the best."""
    ms = np.zeros(len(SORT_CONFIGS), np.float32)
    ctx.check(ctx.lib.trace_autotune_build(ctx.h, mesh.h, key_mode, _ptr(ms)))
    return ms


def autotune_traversal(ctx: Context, bvh: Bvh, d_rays: int, n: int) -> np.ndarray:
    """Times (64, 128, 256 threads) x (one ray per lane, persistent in input order,
    persistent with rays grouped by origin); keeps the best."""
    ms = np.zeros(9, np.float32)
    ctx.check(ctx.lib.trace_autotune_traversal(ctx.h, bvh.h, C.c_void_p(d_rays), n, _ptr(ms)))
    return ms


def autotune_render(ctx: Context, mesh: Mesh, bvh: Bvh, cam, width: int, height: int, spp: int, max_depth: int,
                    miss: int, seed: int) -> np.ndarray:
    ms = np.zeros(9, np.float32)
    c = camera_struct(cam)
    p = render_params(width, height, spp, max_depth, miss, seed)
    ctx.check(ctx.lib.trace_autotune_render(ctx.h, mesh.h, bvh.h, C.byref(c), C.byref(p), _ptr(ms)))
    return ms
