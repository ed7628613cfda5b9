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
kernel and the FP64 CPU oracle timed on a sample.

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

    threads = max(1, min(16, os.cpu_count() or 1))
    tree = pyoracle.PhrTree(pyoracle.Tree(pos))
    probe = 20_000
    t0 = time.perf_counter()
    tree.closest_hit(rays[:probe], threads)
    dt = time.perf_counter() - t0
    n = int(min(len(rays), max(probe, probe * budget_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    tree.closest_hit(rays[:n], threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(n / dt / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} of the 8M incoherent rays, FP64 restatement of BVH.intersected (recursive, "
                  f"stored child order) on the oracle's PHR tree of the same 1M triangles, "
                  f"{threads} threads, {dt:.1f} s",
    }


def load_traffic():
    path = os.path.join(ROOT, "profiles", "closest_hit_incoherent_1M_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--rays", type=int, default=8_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--quick", action="store_true", help="skip the extra measurements")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path would shard rays/pixels across GPUs with a replicated BVH)", file=sys.stderr)
        return 2

    import torch

    from trace_amd import KEY_M30, KEY_REF63, scenes
    from trace_amd.device import Bvh, Context, Mesh

    pos = scenes.random_triangles(args.tris, seed=1)
    rays = scenes.incoherent_rays(args.rays, seed=2)
    ctx = Context(0)
    # A dedicated (default-priority) stream: torch's default stream has handle
    # 0, which the C-ABI reads as "use the context's own stream".  Events are
    # recorded on this same stream, so they bracket the kernels exactly.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)
    mesh = Mesh(ctx, pos)
    lbvh = Bvh.lbvh(ctx, mesh, key_mode=KEY_REF63)
    t0 = time.perf_counter()
    phr = Bvh.phr(ctx, lbvh)
    phr_ms = (time.perf_counter() - t0) * 1e3

    d_rays = torch.from_numpy(rays).cuda()
    n = rays.shape[0]
    d_hit = torch.empty(n, dtype=torch.int32, device="cuda")
    d_t = torch.empty(n, dtype=torch.float32, device="cuda")

    def tracer(bvh, dr, cnt):
        return lambda: bvh.closest_hit_device(dr.data_ptr(), cnt, d_hit.data_ptr(), d_t.data_ptr())

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
            "kernel": "closest_hit_kernel",
            "kernel_ms": round(launch_ms, 4),
            "bytes_per_ray": round(bytes_launch / n, 1),
            "inner_nodes_per_ray": round(float(nodes.mean()), 2),
            "tris_per_ray": round(float(tris.mean()), 2),
        },
        "hit_fraction": round(float((hit >= 0).mean()), 4),
        "phr_build_ms": round(phr_ms, 2),
    }

    if not args.quick:
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
        for name, mode in (("ref63", KEY_REF63), ("m30", KEY_M30)):
            b = Bvh.lbvh(ctx, mesh, key_mode=mode)
            per_b, _ = time_launches(torch, lambda: b.rebuild(mesh, key_mode=mode), max(5, args.steps // 2), 2)
            ctx.synchronize()
            builds[name] = {"ms": round(float(per_b.mean()), 4),
                            "mtris_per_s": round(args.tris / (float(per_b.mean()) * 1e-3) / 1e6, 1)}
            b.close()
        out["lbvh_build"] = builds

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pos, rays)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
