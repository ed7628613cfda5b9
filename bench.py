"""Benchmark of the rendering hot path on one MI355X.

Workload (the build + traverse microbenchmark): 1,000,000 triangles with
centroids ~U[0,1]^3 and vertex offsets of length ~Exp(mean 0.005) (seed 1);
8,000,000 incoherent rays (origins ~U[0,1]^3, directions uniform on the
sphere, seed 2) and 4096x2048 coherent pinhole rays.  Pipeline: LBVH
(63-bit keys) -> PHR refinement (alpha 0.5, delta 6, branching 2) -> closest
hit, as the reference's Scene.Compile does.

A step is one closest-hit launch over the 8M incoherent rays on the PHR tree,
with rays, BVH and outputs resident in HBM.  `value` = incoherent Mrays/s.
The line also carries closest hit on the LBVH, coherent rays, the LBVH build
(63-bit and 30-bit keys) and PHR build times, the roofline of the traversal
kernel, the FP64 CPU oracle timed on a sample, and the other workloads: the
small render (640x360, 16 spp), batched builds and renders (64 meshes, 32
scenes), a 120-frame animation under one HIP graph, and the adjoint/primal
cost of the differentiation workload (200k triangles, 512x512, 32 spp), and
the large configuration (10.4M triangles, PHR tree, 1920x1080, 64 spp, depth
8, Adam steps of inverse rendering).

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s incoherent closest-hit @1M tris; LBVH build ms; adjoint/primal cost"


def note(msg):
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _events(torch, n):
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def time_launches(torch, fn, steps, warmup):
    """Per-launch device times (ms) from events on the current stream, and the
    wall time (ms) of the whole timed region."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = _events(torch, steps), _events(torch, steps)
    t0 = time.perf_counter()
    for k in range(steps):
        s[k].record()
        fn()
        e[k].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    per = np.array([a.elapsed_time(b) for a, b in zip(s, e)])
    return per, wall


def cpu_baseline(pos, rays, budget_s=12.0):
    """FP64 oracle: the reference's recursive traversal in stored child order on
    the same PHR tree (Scene.Compile's default builder), on a bounded sample of
    the same rays, using the host's cores."""
    from oracle import pyoracle

    threads = max(1, os.cpu_count() or 1)
    tree = pyoracle.PhrTree(pyoracle.Tree(pos))
    probe = 20_000
    t0 = time.perf_counter()
    tree.closest_hit(rays[:probe], threads)
    dt = time.perf_counter() - t0
    n = int(min(len(rays), max(probe, probe * budget_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    tree.closest_hit(rays[:n], threads)
    dt = time.perf_counter() - t0
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    return {
        "value": round(n / dt / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "host_cpus": os.cpu_count(),
        "cpu_model": model,
        "kind": "restatement",
        "sample": f"first {n} of the 8M incoherent rays, FP64 C++ restatement (not the Go path) of "
                  f"BVH.intersected (recursive, "
                  f"stored child order) on the oracle's PHR tree of the same 1M triangles, "
                  f"{threads} threads, {dt:.1f} s",
    }


def bench_render_small(torch, ctx, steps):
    """The small plain-method scene: 640x360, 16 spp, depth 5, box + light +
    20k-triangle jittered sphere, 30-bit-key LBVH, default (black) miss."""
    from trace_amd import KEY_M30, _lib, scenes
    from trace_amd.device import Bvh, Mesh, render_device

    sc = scenes.box_scene(seed=1)
    W, H, spp, depth = sc["width"], sc["height"], 16, 5
    mesh = Mesh(ctx, sc["pos"])
    mesh.set_shading(sc["mat_idx"], sc["materials"])
    t0 = time.perf_counter()
    bvh = Bvh.lbvh(ctx, mesh, key_mode=KEY_M30)
    ctx.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    s = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    q = torch.empty_like(s)
    per, _ = time_launches(torch, lambda: render_device(ctx, mesh, bvh, sc["camera"], W, H, spp, depth,
                                                        _lib.MISS_DEFAULT, 1, s.data_ptr(), q.data_ptr()),
                           steps, 1)
    ctx.synchronize()
    ms = float(np.median(per))
    return {"tris": len(sc["pos"]), "image": [W, H], "spp": spp, "max_depth": depth, "render_ms": round(ms, 3),
            "samples_per_s": round(W * H * spp / (ms * 1e-3), 1), "build_wall_ms": round(build_ms, 2)}


def bench_batched(torch, ctx, steps):
    """64 jittered spheres with log-uniform [1k, 500k] triangles: one forest
    build against a loop of per-mesh builds; then 32 of them rendered at
    512x512, 8 spp, depth 4 in one wavefront against a loop of renders."""
    from trace_amd import _lib, scenes
    from trace_amd.device import Bvh, Mesh, render_batched, render_device

    meshes = scenes.batch_meshes(64, 1000, 500_000, seed=11)
    total = sum(len(m) for m in meshes)
    seg = Mesh.segmented(ctx, meshes)
    forest = Bvh.lbvh(ctx, seg)
    per_f, _ = time_launches(torch, lambda: forest.rebuild(seg), steps, 1)
    singles = [Mesh(ctx, m) for m in meshes]
    bvhs = [Bvh.lbvh(ctx, m) for m in singles]

    def loop():
        for b, m in zip(bvhs, singles):
            b.rebuild(m)

    per_l, _ = time_launches(torch, loop, max(3, steps // 2), 1)
    ctx.synchronize()
    for b in bvhs:
        b.close()
    for m in singles:
        m.close()
    out = {"meshes": 64, "tris": total, "forest_build_ms": round(float(np.median(per_f)), 3),
           "per_mesh_loop_ms": round(float(np.median(per_l)), 3)}
    out["speedup"] = round(out["per_mesh_loop_ms"] / out["forest_build_ms"], 2)
    out["forest_mtris_per_s"] = round(total / (out["forest_build_ms"] * 1e-3) / 1e6, 1)

    ns, W, H, spp, depth = 32, 512, 512, 8, 4
    mi = np.zeros(total, np.int32)
    seg.set_shading(mi, np.array([[scenes.MAT_DIFFUSE, 0.7, 0.6, 0.5]], np.float32))
    cams = []
    for m in meshes[:ns]:
        v = m.reshape(-1, 3)
        c = (v.min(0) + v.max(0)) / 2
        r = float((v.max(0) - v.min(0)).max()) / 2
        cams.append(scenes.Camera(1.0, 40.0, tuple(c + np.array([0.3, 0.5, 3.2]) * r), tuple(c), (0, 1, 0)))
    img = torch.empty((ns, H, W, 3), dtype=torch.float32, device="cuda")
    sq = torch.empty_like(img)
    segs = list(range(ns))
    per_b, _ = time_launches(torch, lambda: render_batched(ctx, seg, forest, segs, cams, W, H, spp, depth,
                                                           _lib.MISS_SKY, 3, img.data_ptr(), sq.data_ptr()),
                             max(3, steps // 4), 1)
    views = [forest.segment(i) for i in range(ns)]

    def loop_render():
        for i in range(ns):
            render_device(ctx, seg, views[i], cams[i], W, H, spp, depth, _lib.MISS_SKY, 3 + i, img[i].data_ptr(),
                          sq[i].data_ptr())

    per_r, _ = time_launches(torch, loop_render, max(3, steps // 4), 1)
    ctx.synchronize()
    for v in views:
        v.close()
    b_ms, r_ms = float(np.median(per_b)), float(np.median(per_r))
    samples = ns * W * H * spp
    out["render"] = {"scenes": ns, "image": [W, H], "spp": spp, "max_depth": depth, "batched_ms": round(b_ms, 2),
                     "loop_ms": round(r_ms, 2), "samples_per_s": round(samples / (b_ms * 1e-3), 1),
                     "speedup": round(r_ms / b_ms, 2)}
    forest.close()
    seg.close()
    return out


def bench_animation(torch, ctx, frames=120, tris=200_000, phr=False):
    """120 frames of a jittered sphere (200k triangles; 4k for the
    launch-bound case, where a frame's ~20 kernels each run for a few
    microseconds) under seeded sinusoidal vertex displacement (amplitude
    0.02), rebuilt every frame: all frames recorded in one HIP graph, against
    the same calls issued eagerly.  phr: every frame also refines its LBVH
    into one PHR tree's storage (alpha 0.5, delta 6, branching 2), whose
    eager form reads the level counters back to the host."""
    from trace_amd import scenes
    from trace_amd.device import Bvh, Graph, Mesh

    pos = scenes.displaced_sphere(tris, seed=21, radius=1.0)
    rest = Mesh(ctx, pos)
    mesh = Mesh(ctx, pos)
    bvh = Bvh.lbvh(ctx, mesh)
    tree = None
    if phr:
        tree = Bvh.phr(ctx, bvh)
        tree.phr_rebuild(bvh)  # sizes the tree for the captured rebuilds
    frame = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    def one():
        mesh.animate(rest, 0.02, 60.0, 5, frame.data_ptr())
        bvh.rebuild(mesh)
        if tree is not None:
            tree.phr_rebuild(bvh)

    def eager():
        for _ in range(frames):
            one()

    per_e, _ = time_launches(torch, eager, 3, 1)
    ctx.synchronize()
    graph = Graph.capture(ctx, eager)
    per_g, _ = time_launches(torch, graph.launch, 3, 1)
    ctx.synchronize()
    graph.close()
    e_ms, g_ms = float(np.median(per_e)), float(np.median(per_g))
    out = {"tris": len(pos), "frames": frames, "graph_ms_per_frame": round(g_ms / frames, 4),
           "eager_ms_per_frame": round(e_ms / frames, 4), "graph_speedup": round(e_ms / g_ms, 2),
           "frames_done": int(frame.cpu()[0])}
    if tree is not None:
        out["phr_levels"] = tree.info().phr_levels
        tree.close()
    bvh.close()
    mesh.close()
    rest.close()
    return out


def bench_adjoint(torch, ctx, steps):
    """The differentiation workload: a 200k-triangle scene (box, four jittered
    spheres, area light) at 512x512, 32 spp, depth 5; target rendered from the
    ground-truth albedo (seed 7), current albedo 0.5, L2 image loss.  Times one
    primal render and one backward pass (path replay) of the same paths."""
    from trace_amd import _lib, scenes
    from trace_amd.device import Bvh, Mesh, render_backward, render_device

    d = scenes.adjoint_scene(200_000, seed=7, width=512, height=512)
    W, H, spp, depth = 512, 512, 32, 5
    mesh = Mesh(ctx, d["pos"])
    mesh.set_shading(d["mat_idx"], d["materials"])
    bvh = Bvh.lbvh(ctx, mesh)
    n = len(d["pos"])
    true_alb = torch.from_numpy(d["albedo"]).cuda()
    emi = torch.from_numpy(d["emission"]).cuda()
    alb = torch.full((n, 3), 0.5, dtype=torch.float32, device="cuda")
    s = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    q = torch.empty_like(s)
    torch.cuda.synchronize()
    mesh.bind_params(true_alb.data_ptr(), emi.data_ptr())
    render_device(ctx, mesh, bvh, d["camera"], W, H, spp, depth, _lib.MISS_DEFAULT, 7, s.data_ptr(), q.data_ptr())
    ctx.synchronize()
    target = s / spp
    mesh.bind_params(alb.data_ptr(), emi.data_ptr())

    def primal():
        render_device(ctx, mesh, bvh, d["camera"], W, H, spp, depth, _lib.MISS_DEFAULT, 7, s.data_ptr(),
                      q.data_ptr())

    per_p, _ = time_launches(torch, primal, steps, 1)
    ctx.synchronize()
    g_img = (2.0 / target.numel()) * (s / spp - target)
    loss = float(((s / spp - target) ** 2).mean())
    ga = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    ge = torch.empty_like(ga)
    torch.cuda.synchronize()

    def backward():
        render_backward(ctx, mesh, bvh, d["camera"], W, H, spp, depth, _lib.MISS_DEFAULT, 7, g_img.data_ptr(),
                        ga.data_ptr(), ge.data_ptr())

    per_b, _ = time_launches(torch, backward, steps, 1)
    ctx.synchronize()
    p_ms, b_ms = float(np.median(per_p)), float(np.median(per_b))
    # the same pair with the path summaries kept by the primal (16 B per path):
    # the backward then traces only the replay sweep
    from trace_amd.device import render_backward_kept, render_keep
    keep = torch.empty((spp * H * W * 4,), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def primal_keep():
        render_keep(ctx, mesh, bvh, d["camera"], W, H, spp, depth, _lib.MISS_DEFAULT, 7, s.data_ptr(), q.data_ptr(),
                    keep.data_ptr())

    def backward_kept():
        render_backward_kept(ctx, mesh, bvh, d["camera"], W, H, spp, depth, _lib.MISS_DEFAULT, 7, keep.data_ptr(),
                             g_img.data_ptr(), ga.data_ptr(), ge.data_ptr())

    per_pk, _ = time_launches(torch, primal_keep, steps, 1)
    per_bk, _ = time_launches(torch, backward_kept, steps, 1)
    ctx.synchronize()
    pk_ms, bk_ms = float(np.median(per_pk)), float(np.median(per_bk))
    out = {"tris": n, "image": [W, H], "spp": spp, "max_depth": depth, "primal_ms": round(p_ms, 3),
           "backward_ms": round(b_ms, 3), "adjoint_over_primal": round(b_ms / p_ms, 3), "loss": loss,
           "grad_albedo_norm": float(torch.linalg.vector_norm(ga).cpu()),
           "grad_emission_norm": float(torch.linalg.vector_norm(ge).cpu()),
           "kept": {"primal_keep_ms": round(pk_ms, 3), "backward_kept_ms": round(bk_ms, 3),
                    "backward_kept_over_primal": round(bk_ms / p_ms, 3),
                    "step_ms": round(pk_ms + bk_ms, 3), "step_ms_without_keep": round(p_ms + b_ms, 3)}}
    mesh.bind_params(None, None)
    bvh.close()
    mesh.close()
    return out


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "closest_hit_incoherent_1M_traffic.json")
COPY_BYTES = 1 << 30  # reference copy kernel: reads and writes 1 GiB


def bench_heatmap(torch, ctx, bvh, steps):
    """HeatMapRenderer.RenderToBuffer on the 1M-triangle PHR tree: 1920x1080
    camera rays from outside the unit cube (the coherent-ray camera),
    threshold 50; FP64 reference-order walk, one ray per lane."""
    import ctypes

    from trace_amd import scenes
    from trace_amd.device import camera_struct

    W, H, th = 1920, 1080, 50
    cam = scenes.camera_cfg2(W, H)
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    cnt = torch.empty((H, W), dtype=torch.int32, device="cuda")
    c = camera_struct(cam)
    torch.cuda.synchronize()

    def run():
        ctx.check(ctx.lib.trace_render_heatmap_device(ctx.h, bvh.h, ctypes.byref(c), W, H, th, rgb.data_ptr(),
                                                      cnt.data_ptr()))

    per, _ = time_launches(torch, run, steps, 1)
    ms = float(np.median(per))
    counts = cnt.cpu().numpy()
    return {"image": [W, H], "threshold": th, "ms": round(ms, 3), "mrays_s": round(W * H / (ms * 1e3), 1),
            "mean_steps": round(float(counts.mean()), 2), "max_steps": int(counts.max())}


def bench_scene_compile(torch, ctx, steps):
    """Scene.collectTracables on the device for the large configuration's scene
    graph: the 2048x2048 terrain mesh plus 5,000 placements (scale and
    translate) of one 400-triangle sphere mesh -> 10,380,418 triangles with
    shading; then the LBVH of the result."""
    from trace_amd import scenes
    from trace_amd.device import Bvh, Scene

    meshes, mm, parent, node_mesh, mats = scenes.terrain_scene_graph()
    materials = np.array([[scenes.MAT_DIFFUSE, 0.5, 0.5, 0.5], [scenes.MAT_DIFFUSE, 0.8, 0.8, 0.8]], np.float32)
    t0 = time.perf_counter()
    sc = Scene(ctx, meshes, mm, materials)
    upload_ms = (time.perf_counter() - t0) * 1e3
    mesh = sc.compile(parent, node_mesh, mats)
    torch.cuda.synchronize()
    per, _ = time_launches(torch, lambda: sc.compile(parent, node_mesh, mats, mesh=mesh), steps, 1)
    bvh = Bvh.lbvh(ctx, mesh)
    ctx.synchronize()
    # the same scene as flattened on the host by the generator (FP32-rounded there)
    ref = scenes.terrain_scene()["pos"]
    err = float(np.abs(mesh.export() - ref).max())
    out = {"tris": int(mesh.ntri), "nodes": int(len(parent)), "source_tris": int(sum(len(m) for m in meshes)),
           "upload_ms": round(upload_ms, 1), "compile_ms": round(float(np.median(per)), 3),
           "max_abs_diff_vs_host_flatten": err, "lbvh_n_leaves": int(bvh.info().n_leaves)}
    bvh.close()
    mesh.close()
    sc.close()
    return out


def build_roofline(ms, ntri, key_mode_bits, sort_bits):
    """HBM roofline of one LBVH build: the algorithmic bytes per triangle
    36 (positions) + (kb + 4)(1 + 2P) (keys and values written once, then read
    and written per radix pass) + 64 (two 32-byte nodes) + 4 (permutation) +
    24 (the refit re-reads the triangle box), with kb = 8-byte keys for 63-bit
    codes, 4-byte for 30-bit ones, and P = ceil(key bits / digit bits)."""
    kb = 8 if key_mode_bits > 32 else 4
    passes = -(-key_mode_bits // sort_bits)
    per_tri = 36 + (kb + 4) * (1 + 2 * passes) + 64 + 4 + 24
    achieved = per_tri * ntri / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "bytes_per_tri": per_tri, "passes": passes, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)}


def bench_instancing(torch, ctx, steps):
    """The large configuration's 5,000 sphere placements as instances of one
    400-triangle tree under a top level over the instances (plus the terrain
    as one more instance), against the flattened 10.38M-triangle scene: build
    time, closest hit on the 1920x1080 camera rays, device memory."""
    from trace_amd import scenes
    from trace_amd.device import Bvh, Mesh

    meshes, _, _, node_mesh, mats = scenes.terrain_scene_graph()
    terrain = meshes[0].astype(np.float32).astype(np.float64)
    unit = meshes[1].astype(np.float32).astype(np.float64)
    xforms = mats.reshape(-1, 4, 4)[:, :3, :].reshape(-1, 12)
    base = np.concatenate([[0], len(terrain) + len(unit) * np.arange(len(node_mesh) - 1)]).astype(np.int64)
    flat = np.concatenate([terrain] + [(unit.reshape(-1, 3) @ x.reshape(3, 4)[:, :3].T + x.reshape(3, 4)[:, 3])
                                       .reshape(-1, 9) for x in xforms[1:]])
    d = scenes.terrain_scene()
    rays = torch.from_numpy(scenes.camera_rays(d["camera"], d["width"], d["height"]).astype(np.float32)).cuda()
    n = rays.shape[0]
    hit = torch.empty(n, dtype=torch.int32, device="cuda")
    t = torch.empty(n, dtype=torch.float32, device="cuda")
    fmesh = Mesh(ctx, flat)
    fl = Bvh.lbvh(ctx, fmesh)
    per_fb, _ = time_launches(torch, lambda: fl.rebuild(fmesh), 3, 1)
    blas = [Bvh.lbvh(ctx, Mesh(ctx, terrain)), Bvh.lbvh(ctx, Mesh(ctx, unit))]
    ctx.synchronize()
    t0 = time.perf_counter()
    inst = Bvh.instances(ctx, blas, node_mesh, xforms, base)
    tlas_ms = (time.perf_counter() - t0) * 1e3
    tmesh = Mesh(ctx, terrain)
    per_tb, _ = time_launches(torch, lambda: blas[0].rebuild(tmesh), 3, 1)
    per_f, _ = time_launches(torch, lambda: fl.closest_hit_device(rays.data_ptr(), n, hit.data_ptr(), t.data_ptr()),
                             steps, 2)
    per_i, _ = time_launches(torch, lambda: inst.closest_hit_device(rays.data_ptr(), n, hit.data_ptr(), t.data_ptr()),
                             steps, 2)
    ctx.synchronize()
    fb = fl.device_bytes()
    ib = inst.device_bytes() + sum(b.device_bytes() for b in blas)
    out = {"instances": int(len(node_mesh)), "flattened_tris": int(len(flat)), "rays": int(n),
           "flattened_build_ms": round(float(np.median(per_fb)), 3),
           "terrain_build_ms": round(float(np.median(per_tb)), 3), "tlas_build_wall_ms": round(tlas_ms, 2),
           "flattened_mrays_s": round(n / (float(np.median(per_f)) * 1e3), 1),
           "instanced_mrays_s": round(n / (float(np.median(per_i)) * 1e3), 1),
           "flattened_tree_bytes": fb, "instanced_tree_bytes": ib, "bytes_saved": fb - ib}
    for b in (inst, fl, *blas):
        b.close()
    return out


def bench_large(torch, ctx, steps=100):
    """The large configuration: 10.38M triangles (2048x2048 fBm terrain + 5000
    400-triangle spheres), PHR tree, 1920x1080, 64 spp, depth 8, sky miss;
    inverse rendering of the per-triangle terrain albedo with Adam (lr 1e-2)
    from 0.5 towards a target rendered from the ground truth, L2 image loss.
    Times the builds, one primal render and `steps` optimisation steps
    (primal + backward + Adam)."""
    from trace_amd import _lib, scenes
    from trace_amd.device import (Bvh, Mesh, adam_step, render_backward_kept, render_device, render_keep,
                                  render_segments)

    d = scenes.terrain_scene()
    n, nt = len(d["pos"]), d["n_terrain"]
    W, H, spp, depth, seed = d["width"], d["height"], 64, 8, 9
    mesh = Mesh(ctx, d["pos"])
    mesh.set_shading(d["mat_idx"], d["materials"])
    from trace_amd.device import SORT_CONFIGS, autotune_build
    sort_ms = autotune_build(ctx, mesh)  # the key sort tuned on this scene's own build
    lbvh = Bvh.lbvh(ctx, mesh)
    ev = _events(torch, 4)
    ev[0].record()
    lbvh.rebuild(mesh)
    ev[1].record()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phr = Bvh.phr(ctx, lbvh)
    ctx.synchronize()
    phr_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    phr.phr_rebuild(lbvh)
    ctx.synchronize()
    phr_rebuild_ms = (time.perf_counter() - t0) * 1e3
    true_alb = torch.from_numpy(d["albedo"]).cuda()
    emi = torch.from_numpy(d["emission"]).cuda()
    alb = true_alb.clone()
    alb[:nt] = 0.5
    m = torch.zeros_like(alb)
    v = torch.zeros_like(alb)
    img = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    sq = torch.empty_like(img)
    ga = torch.empty_like(alb)
    ge = torch.empty_like(alb)
    torch.cuda.synchronize()
    mesh.bind_params(true_alb.data_ptr(), emi.data_ptr())
    ev[2].record()
    render_device(ctx, mesh, phr, d["camera"], W, H, spp, depth, _lib.MISS_SKY, seed, img.data_ptr(), sq.data_ptr())
    ev[3].record()
    torch.cuda.synchronize()
    render_rays = render_segments(ctx)
    # byte accounting of the same render: B_seg = 32 (ray in) + 8 (hit out) +
    # 64 per 4-wide node visited + 36 per primitive tested + 120 of path state
    # (hit record, shading data, throughput, radiance, new ray, queue index)
    from trace_amd.device import render_counts
    segs, visits, tests = render_counts(ctx, mesh, phr, d["camera"], W, H, spp, depth, _lib.MISS_SKY, seed)
    render_roof = {"bound": "hbm", "segments": segs, "nodes_per_segment": round(visits / segs, 2),
                   "tris_per_segment": round(tests / segs, 2)}
    target = img / spp
    mesh.bind_params(alb.data_ptr(), emi.data_ptr())
    # the primal keeps every path's summary (16 B per path, 2.1 GB here) so
    # that the backward traces only its replay sweep
    keep = torch.empty((spp * H * W * 4,), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    losses, step_ms, step_rays = [], [], []
    for k in range(1, steps + 1):
        if k == 1 or k % 10 == 0:
            note(f"large: Adam step {k} of {steps}")
        e0, e1, e2 = _events(torch, 3)
        e0.record()
        render_keep(ctx, mesh, phr, d["camera"], W, H, spp, depth, _lib.MISS_SKY, seed, img.data_ptr(),
                    sq.data_ptr(), keep.data_ptr())
        e1.record()
        diff = img / spp - target
        g_img = (2.0 / diff.numel()) * diff
        render_backward_kept(ctx, mesh, phr, d["camera"], W, H, spp, depth, _lib.MISS_SKY, seed, keep.data_ptr(),
                             g_img.data_ptr(), ga.data_ptr(), ge.data_ptr())
        adam_step(ctx, alb.data_ptr(), ga.data_ptr(), m.data_ptr(), v.data_ptr(), 3 * nt, 1e-2, k)
        e2.record()
        torch.cuda.synchronize()
        # rays of the primal render (the same paths as the target's) and of the replay sweep
        step_rays.append(render_rays + render_segments(ctx))
        losses.append(float((diff * diff).mean()))
        step_ms.append(e0.elapsed_time(e2))
    ctx.synchronize()
    render_ms = ev[2].elapsed_time(ev[3])
    build_ms = ev[0].elapsed_time(ev[1])
    b_seg = 32 + 8 + 64 * visits / segs + 36 * tests / segs + 120
    ach = b_seg * segs / (render_ms * 1e-3) / 1e9
    render_roof.update({"bytes_per_segment": round(b_seg, 1), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)})
    best = SORT_CONFIGS[int(np.argmin(sort_ms))]
    out = {"tris": n, "terrain_tris": nt, "image": [W, H], "spp": spp, "max_depth": depth,
           "lbvh_build_ms": round(build_ms, 3),
           "lbvh_build_roofline": build_roofline(build_ms, n, 63, best[0]),
           "lbvh_sort_ms_by_config": {f"{c[0]}x{c[1]}": round(float(m), 3) for c, m in zip(SORT_CONFIGS, sort_ms)},
           "phr_build_ms": round(phr_ms, 2),
           "phr_rebuild_ms": round(phr_rebuild_ms, 2),
           "render_ms": round(render_ms, 2), "render_rays": render_rays, "render_roofline": render_roof,
           "render_mrays_s": round(render_rays / (render_ms * 1e3), 1),
           "samples_per_s": round(W * H * spp / (render_ms * 1e-3), 1),
           "adam_steps": steps, "loop_s": round(sum(step_ms) * 1e-3, 3),
           "step_ms": [round(x, 2) for x in step_ms], "step_rays": step_rays,
           "step_mrays_s": [round(r / (t * 1e3), 1) for r, t in zip(step_rays, step_ms)],
           "loss": [float(f"{x:.6e}") for x in losses]}
    mesh.bind_params(None, None)
    phr.close()
    lbvh.close()
    mesh.close()
    return out


def load_traffic():
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def pmc_probe(torch, ctx, phr, d_rays, n, d_hit, d_t):
    """Workload for the two counter runs (FETCH_SIZE, WRITE_SIZE): a plain
    copy of COPY_BYTES (torch.mul by 1 into a second buffer, read once and
    written once), then PROBE_CALLS closest-hit calls of the timed workload."""
    from trace_amd.device import get_tune, set_tune

    # counter collection serialises and slows every dispatch, so the
    # autotuner's pick in this process may differ from the timed run's: the
    # probe uses the mode the timed run picks (persistent lanes, rays grouped
    # by origin) with the workgroup size tuned here
    set_tune(ctx, trav_threads=int(get_tune(ctx).trav_threads), trav_persistent=True, trav_order=True)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    for _ in range(PROBE_CALLS):
        phr.closest_hit_device(d_rays.data_ptr(), n, d_hit.data_ptr(), d_t.data_ptr())
    ctx.synchronize()


def pmc_probe_build(torch, ctx):
    """Workload for the build's counter runs: the same copy of COPY_BYTES, then
    the 10.38M-triangle terrain LBVH (63-bit keys, default key sort) built once
    and rebuilt twice; the last rebuild is the one a summary reads."""
    from trace_amd import scenes
    from trace_amd.device import Bvh, Mesh

    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst
    mesh = Mesh(ctx, scenes.terrain_scene()["pos"])
    bvh = Bvh.lbvh(ctx, mesh)
    for _ in range(2):
        bvh.rebuild(mesh)
    ctx.synchronize()
    bvh.close()
    mesh.close()


def _counter_rows(path):
    import csv

    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
            val = r.get("Counter_Value") or r.get("Counter-Value")
            cname = r.get("Counter_Name") or r.get("Counter-Name") or ""
            if val is None:
                continue
            did = r.get("Dispatch_Id") or r.get("Dispatch-Id") or len(rows)
            rows.append((int(did), name, cname, float(val)))
    rows.sort()
    return [(k, c, v) for _, k, c, v in rows]


# The kernels of one closest-hit call (traverse.hip): the work-counter fill,
# the origin-order passes (probe, cell counts, their scan, the cell-major
# copy), the persistent traversal, then the FP64 re-solve in input order or,
# for an ordered batch, in traced order followed by the gather back.
PROBE_CALLS = 3
CALL_KERNELS = ("fill_kernel", "traverse::cell_probe_kernel", "traverse::cell_count_kernel", "scan::tile_sums",
                "scan::scan_sums", "scan::tile_scan", "traverse::cell_scatter_kernel", "traverse::closest_hit",
                "traverse::refine_kernel", "traverse::refine_traced_kernel", "traverse::unpermute_kernel")


def traffic_from(fetch_csv, write_csv, hitmiss_csv=None):
    """HBM bytes per closest-hit call from the two counter runs, each scaled
    so that the copy kernel reads / writes exactly COPY_BYTES: every dispatch
    after the copy (the probe's PROBE_CALLS calls, CALL_KERNELS) summed and
    divided by PROBE_CALLS; with a third run of TCC_HIT_sum and TCC_MISS_sum,
    the traversal kernel's L2 hit rate."""
    out = {}
    for label, path in (("fetch", fetch_csv), ("write", write_csv)):
        rows = _counter_rows(path)
        copies = [i for i, (k, c, v) in enumerate(rows) if "elementwise" in k.lower() or "mul" in k.lower()]
        if not copies:
            raise SystemExit(f"{path}: copy kernel dispatch not found")
        copy_kib = rows[copies[-1]][2]
        per = {}
        for k, c, v in rows[copies[-1] + 1:]:
            if k.startswith("__amd_rocclr"):  # the runtime's own copies (error word, results)
                continue
            name = next((m for m in CALL_KERNELS if m in k), None)
            if name is None:
                raise SystemExit(f"{path}: unexpected dispatch {k} after the copy")
            per[name] = per.get(name, 0.0) + v / PROBE_CALLS
        if "traverse::closest_hit" not in per:
            raise SystemExit(f"{path}: closest_hit dispatches not found")
        scale = (COPY_BYTES / 1024.0) / copy_kib
        kib = sum(per.values())
        out[label] = {"copy_kib": copy_kib, "scale": scale, "kib_raw_by_kernel": per,
                      "kernel_bytes": kib * 1024.0 * scale,
                      "traversal_bytes": per["traverse::closest_hit"] * 1024.0 * scale}
    total = out["fetch"]["kernel_bytes"] + out["write"]["kernel_bytes"]
    res = {"hbm_bytes_per_launch": round(total), "fetch_bytes": round(out["fetch"]["kernel_bytes"]),
           "write_bytes": round(out["write"]["kernel_bytes"]),
           "traversal_kernel_bytes": round(out["fetch"]["traversal_bytes"] + out["write"]["traversal_bytes"]),
           "fetch_scale": round(out["fetch"]["scale"], 4),
           "write_scale": round(out["write"]["scale"], 4), "copy_bytes": COPY_BYTES,
           "raw": out}
    if hitmiss_csv:
        rows = [(k, c, v) for k, c, v in _counter_rows(hitmiss_csv) if "traverse::closest_hit" in k]
        hits = [v for k, c, v in rows if c.startswith("TCC_HIT")][-PROBE_CALLS:]
        miss = [v for k, c, v in rows if c.startswith("TCC_MISS")][-PROBE_CALLS:]
        if hits and miss:
            res["l2_hit_rate"] = round(float(np.sum(hits) / (np.sum(hits) + np.sum(miss))), 4)
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--rays", type=int, default=8_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--quick", action="store_true", help="skip the extra measurements")
    ap.add_argument("--no-large", action="store_true", help="skip the 10.4M-triangle inverse-rendering leg")
    ap.add_argument("--adam-steps", type=int, default=100,
                    help="inverse-rendering steps of the 10.4M-triangle leg (the configuration's 100)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run only the counter workload (copy kernel + closest hit) and exit")
    ap.add_argument("--pmc-probe-build", action="store_true",
                    help="run only the build's counter workload (copy kernel + 10.4M-triangle LBVH) and exit")
    ap.add_argument("--traffic-from", nargs="+", metavar="CSV",
                    help="FETCH_CSV WRITE_CSV [HITMISS_CSV]: turn rocprofv3 counter CSVs into the traffic "
                         "file and exit")
    args = ap.parse_args()
    if args.traffic_from:
        return traffic_from(*args.traffic_from)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path would shard rays/pixels across GPUs with a replicated BVH)", file=sys.stderr)
        return 2

    import torch

    from trace_amd import KEY_M30, KEY_REF63, scenes
    from trace_amd.device import Bvh, Context, Mesh, autotune_traversal, get_tune

    ctx = Context(0)
    # A dedicated (default-priority) stream: torch's default stream has handle
    # 0, which the C-ABI reads as "use the context's own stream".  Events are
    # recorded on this same stream, so they bracket the kernels exactly.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)
    if args.pmc_probe_build:
        pmc_probe_build(torch, ctx)
        ctx.close()
        return 0
    pos = scenes.random_triangles(args.tris, seed=1)
    rays = scenes.incoherent_rays(args.rays, seed=2)
    mesh = Mesh(ctx, pos)
    lbvh = Bvh.lbvh(ctx, mesh, key_mode=KEY_REF63)
    t0 = time.perf_counter()
    phr = Bvh.phr(ctx, lbvh)
    phr_ms = (time.perf_counter() - t0) * 1e3
    # steady state: rebuilt into the same storage (no allocation)
    t0 = time.perf_counter()
    for _ in range(3):
        phr.phr_rebuild(lbvh)
    ctx.synchronize()
    phr_rebuild_ms = (time.perf_counter() - t0) * 1e3 / 3

    d_rays = torch.from_numpy(rays).cuda()
    n = rays.shape[0]
    d_hit = torch.empty(n, dtype=torch.int32, device="cuda")
    d_t = torch.empty(n, dtype=torch.float32, device="cuda")

    def tracer(bvh, dr, cnt):
        return lambda: bvh.closest_hit_device(dr.data_ptr(), cnt, d_hit.data_ptr(), d_t.data_ptr())

    # the autotuner times each workgroup size on these rays and keeps the fastest
    tune_ms = autotune_traversal(ctx, phr, d_rays.data_ptr(), n)
    tune = get_tune(ctx)
    trav_threads = int(tune.trav_threads)
    trav_persistent = bool(tune.trav_persistent)
    trav_order = trav_persistent and bool(tune.trav_order)
    if args.pmc_probe:
        pmc_probe(torch, ctx, phr, d_rays, n, d_hit, d_t)
        ctx.close()
        return 0

    note(f"trees built, autotuned ({trav_threads} threads, persistent={trav_persistent}); timing closest hit")
    per, wall = time_launches(torch, tracer(phr, d_rays, n), args.steps, args.warmup)
    ctx.synchronize()
    ms_per_step = wall / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    launch_ms = float(per.mean())

    # algorithmic bytes per launch from a counting-mode run of the same traversal
    nodes, tris = phr.counts(rays)
    hit = d_hit.cpu().numpy()
    bytes_launch = (n * (32 + 8) + 4 * int((hit >= 0).sum()) + 64 * int(nodes.astype(np.int64).sum())
                    + 36 * int(tris.astype(np.int64).sum()))
    achieved = bytes_launch / (launch_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "build+traverse microbench: 1M random triangles (seed 1), 8M incoherent rays "
                        "(seed 2), LBVH (63-bit keys) -> PHR (0.5, 6, 2) -> closest hit",
            "tris": args.tris,
            "rays": n,
            "tmin": 0.001,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(),
            # the persistent traversal hands its winners to a short FP64 re-solve pass,
            # and an origin-ordered batch is copied cell-major first and its results
            # gathered back after; kernel_ms brackets all of it (HIP events around the call)
            "kernel": ("cell_probe + cell_count + scan + cell_scatter + closest_hit_persistent + refine_traced"
                       " + unpermute" if trav_order else
                       "closest_hit_persistent + refine_kernel" if trav_persistent else "closest_hit_kernel"),
            "kernel_ms": round(launch_ms, 4),
            "bytes_per_ray": round(bytes_launch / n, 1),
            "inner_nodes_per_ray": round(float(nodes.mean()), 2),
            "tris_per_ray": round(float(tris.mean()), 2),
        },
        "hit_fraction": round(float((hit >= 0).mean()), 4),
        "autotune": {"trav_threads": trav_threads, "trav_persistent": trav_persistent, "trav_order": trav_order,
                     "ms_one_ray_per_lane": {str(t): round(float(m), 4) for t, m in zip((64, 128, 256), tune_ms[:3])},
                     "ms_persistent": {str(t): round(float(m), 4) for t, m in zip((64, 128, 256), tune_ms[3:6])},
                     "ms_persistent_origin_order": {str(t): round(float(m), 4)
                                                    for t, m in zip((64, 128, 256), tune_ms[6:9])}},
        "phr_build_ms": round(phr_ms, 2),
        "phr_rebuild_ms": round(phr_rebuild_ms, 2),
    }

    if not args.quick:
        # the host-buffer entry point (rays in and hits out over PCIe, device
        # staging allocated per call), for comparison with the resident rate
        phr.closest_hit(rays[:1000])
        t0 = time.perf_counter()
        phr.closest_hit(rays)
        out["host_buffers_mrays_s"] = round(n / (time.perf_counter() - t0) / 1e6, 2)
        per_l, _ = time_launches(torch, tracer(lbvh, d_rays, n), max(5, args.steps // 2), 2)
        out["lbvh_incoherent_mrays_s"] = round(n / (float(per_l.mean()) * 1e-3) / 1e6, 2)
        coh = scenes.coherent_rays_cfg2(4096, 2048)
        d_coh = torch.from_numpy(coh).cuda()
        m = coh.shape[0]
        d_hit2 = torch.empty(m, dtype=torch.int32, device="cuda")
        d_t2 = torch.empty(m, dtype=torch.float32, device="cuda")

        def trace_coh(bvh):
            return lambda: bvh.closest_hit_device(d_coh.data_ptr(), m, d_hit2.data_ptr(), d_t2.data_ptr())

        per_c, _ = time_launches(torch, trace_coh(phr), max(5, args.steps // 2), 2)
        out["coherent_mrays_s"] = round(m / (float(per_c.mean()) * 1e-3) / 1e6, 2)
        per_c, _ = time_launches(torch, trace_coh(lbvh), max(5, args.steps // 2), 2)
        out["lbvh_coherent_mrays_s"] = round(m / (float(per_c.mean()) * 1e-3) / 1e6, 2)
        builds = {}
        from trace_amd.device import SORT_CONFIGS, autotune_build
        for name, mode in (("ref63", KEY_REF63), ("m30", KEY_M30)):
            # the autotuner times the key sort's digit width x tile size on this build
            sort_ms = autotune_build(ctx, mesh, mode)
            sort_cfg = SORT_CONFIGS[int(np.argmin(sort_ms))]
            b = Bvh.lbvh(ctx, mesh, key_mode=mode)
            per_b, _ = time_launches(torch, lambda: b.rebuild(mesh, key_mode=mode), max(5, args.steps // 2), 2)
            ctx.synchronize()
            builds[name] = {"ms": round(float(per_b.mean()), 4),
                            "mtris_per_s": round(args.tris / (float(per_b.mean()) * 1e-3) / 1e6, 1),
                            "roofline": build_roofline(float(per_b.mean()), args.tris,
                                                       63 if mode == KEY_REF63 else 30, sort_cfg[0]),
                            "sort": {"bits": sort_cfg[0], "keys_per_thread": sort_cfg[1],
                                     "ms_by_config": {f"{c[0]}x{c[1]}": round(float(m), 4)
                                                      for c, m in zip(SORT_CONFIGS, sort_ms)}}}
            b.close()
        out["lbvh_build"] = builds
        # the reference benchmark's branching factor 4 (benchmark_test.go:22)
        walls = []
        for _ in range(3):
            t0 = time.perf_counter()
            p4 = Bvh.phr(ctx, lbvh, branching=4)
            ctx.synchronize()
            walls.append((time.perf_counter() - t0) * 1e3)
        out["phr_bf4"] = {"build_ms": round(float(np.median(walls)), 2), "cost": round(p4.cost(), 4),
                          "cost_bf2": round(phr.cost(), 4)}
        per_4, _ = time_launches(torch, tracer(p4, d_rays, n), max(5, args.steps // 2), 2)
        out["phr_bf4"]["incoherent_mrays_s"] = round(n / (float(per_4.mean()) * 1e-3) / 1e6, 2)
        p4.close()
        note("PHR grid optimiser (16 BuildWithCost + Cost)")
        from trace_amd.device import phr_optimize_grid
        t0 = time.perf_counter()
        ga, gd, gsah, gbuild = phr_optimize_grid(ctx, lbvh, width=1920, height=1080)
        out["phr_grid"] = {"ms": round((time.perf_counter() - t0) * 1e3, 1), "alpha": ga, "delta": gd,
                           "builds": int(len(gsah)), "sah_min": round(float(gsah.min()), 4),
                           "build_cost_min": int(gbuild.min())}
        note("PHR Bayesian optimiser (2 normaliser builds + 20 evaluations)")
        from trace_amd.device import phr_optimize_bayes
        t0 = time.perf_counter()
        ba, bd, btr = phr_optimize_bayes(ctx, lbvh, width=1920, height=1080, rounds=20, random_rounds=5, seed=1)
        out["phr_bayes"] = {"ms": round((time.perf_counter() - t0) * 1e3, 1), "alpha": round(ba, 5),
                            "delta": round(bd, 5), "builds": int(len(btr)),
                            "objective_min": round(float(np.nanmin(btr[2:, 4])), 6)}
        note("heat map (traversal steps in the reference's walk order, FP64)")
        out["heatmap"] = bench_heatmap(torch, ctx, phr, max(3, args.steps // 4))
        note("traversal and builds timed; small render")
        out["render_small"] = bench_render_small(torch, ctx, max(3, args.steps // 4))
        note("batched builds and renders")
        out["batched"] = bench_batched(torch, ctx, max(3, args.steps // 4))
        note("animation")
        out["animation"] = bench_animation(torch, ctx)
        out["animation_small"] = bench_animation(torch, ctx, tris=4_000)
        out["animation_phr"] = bench_animation(torch, ctx, frames=30, phr=True)
        out["animation_phr_small"] = bench_animation(torch, ctx, frames=30, tris=4_000, phr=True)
        note("adjoint")
        out["adjoint"] = bench_adjoint(torch, ctx, max(3, args.steps // 4))
        out["adjoint_over_primal"] = out["adjoint"]["adjoint_over_primal"]
        # the backward from the summaries its primal kept (replay sweep only)
        out["adjoint_over_primal_kept"] = out["adjoint"]["kept"]["backward_kept_over_primal"]
        if not args.no_large:
            note("large scene (10.4M triangles) and inverse-rendering steps")
            out["large"] = bench_large(torch, ctx, args.adam_steps)
            note("large scene instanced: one sphere tree placed 5000 times")
            out["instancing"] = bench_instancing(torch, ctx, max(3, args.steps // 4))
            note("large scene compiled on the device from its scene graph")
            out["scene_compile"] = bench_scene_compile(torch, ctx, max(3, args.steps // 4))

    if not args.no_cpu:
        note("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(pos, rays)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
