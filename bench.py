#!/usr/bin/env python3
"""Benchmark: closest-hit + sun-shadow throughput on the 1.31M-triangle body.

Workload (one "step"): the synthetic small-body mesh -- icosphere L8
(1,310,720 triangles, R = 2 km, 6-octave value-noise displacement 20%) --
seen by a 2048 x 2048 pinhole camera at a seeded 50-200 km range (FP64 pose,
origin rebased in FP64 before the FP32 traversal), plus one any-hit shadow
ray per primary hit towards a seeded sun direction.  The BVH is built once
(cached body-fixed BVH) and all inputs are resident in HBM; the timed region
contains only the device work of the pass.

Prints ONE JSON line (rank 0).  `value` = (primary + shadow rays) / second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "ceres-raytracer_amd")]

import numpy as np  # noqa: E402

# Reference traversal statistics for this exact workload, measured once with
# the reference CPU tracer (SweepSAH + PRO + NLO BVH, SingleRayTraverser
# Statistics) by tests/golden/make_golden.py; used for the algorithmic bytes.
STATS_PATH = os.path.join(ROOT, "tests", "golden", "c2_reference_stats.json")
# Measured HBM traffic of the dominant kernel at this workload: rocprofv3
# FETCH_SIZE / WRITE_SIZE per dispatch, scaled by the same counters read on a
# plain copy of known size (bench.py --copy-calibration).
TRAFFIC_PATH = os.path.join(ROOT, "profiles", "k_primary_sun_icosphere_L8_2048_traffic.json")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(stats):
    """Per-ray algorithmic HBM bytes (reference BVH, FP32 layout):
    closest: 32 (ray) + 16 (hit) + 64 * steps + 36 * prims
    any-hit: 32 + 4 + 64 * steps + 36 * prims."""
    c = stats["closest"]
    a = stats["shadow"]
    bc = 32 + 16 + 64 * c["traversal_steps_mean"] + 36 * c["intersections_mean"]
    ba = 32 + 4 + 64 * a["traversal_steps_mean"] + 36 * a["intersections_mean"]
    return bc, ba


def cpu_baseline(mesh, cam, sun, budget_s=15.0):
    """Time a bounded sample of the same workload on the host cores.

    Preferred: the reference tracer itself (oracle/_ref, built from the
    reference sources); otherwise the FP64 oracle restatement."""
    import oracle
    ref = None
    try:
        from oracle import ref as refmod
        if refmod.available():
            ref = refmod
    except Exception as e:  # pragma: no cover
        log("reference tracer unavailable: %r" % (e,))
    o, d = oracle.camera_rays(cam)
    rng = np.random.default_rng(123)
    if ref is not None:
        g = ref.Group.from_arrays(mesh.vertices, mesh.indices)
        kind, threads = "reference", ref.num_threads()
        trace = lambda oo, dd: g.trace(oo, dd)
    else:
        s = oracle.Scene(mesh.vertices, mesh.indices)
        kind, threads = "port", oracle.num_threads()
        trace = lambda oo, dd: s.trace(oo, dd)
    n = 1 << 16
    total_rays = 0
    t_sum = 0.0
    while t_sum < budget_s and total_rays < o.shape[0]:
        sel = rng.integers(0, o.shape[0], size=n)
        t0 = time.perf_counter()
        trace(o[sel], d[sel])
        t_sum += time.perf_counter() - t0
        total_rays += n
        n = min(n * 2, 1 << 20)
    return {"value": total_rays / t_sum / 1e6, "unit": "Mrays/s", "cores": int(threads), "kind": kind,
            "sample": "%d primary rays drawn uniformly from the 2048^2 camera of the same workload, closest hit, "
                      "cached BVH (build not timed), %.1f s" % (total_rays, t_sum)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--resolution", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by image tiles with no exchange); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from crt import capi, scenes

    if args.copy_calibration:
        n = 1 << 28   # 2^28 float32 = 1 GiB read + 1 GiB written per copy
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}))
        return

    log("generating icosphere L8")
    mesh = scenes.displaced_icosphere(8, 2000.0, 0.2, 6, 0)
    cam = scenes.small_body_camera(mesh, resolution=args.resolution, seed=1)
    sun = scenes.sun_direction(2)
    log("building BVH on GPU")
    scene = capi.Scene(mesh.vertices, mesh.indices)
    info = scene.info()
    log("BVH: %d wide nodes, %d leaves, depth %d, build %.2f ms, SAH %.2f" % (
        info["n_wide_nodes"], info["n_leaves"], info["max_depth"], info["build_ms"], info["sah_cost"]))
    H = W = args.resolution
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), dtype=torch.float64, device=dev)
    tri = torch.empty((H, W), dtype=torch.int32, device=dev)
    lit = torch.empty((H, W), dtype=torch.uint8, device=dev)
    outs = {"depth": depth.data_ptr(), "tri": tri.data_ptr(), "lit": lit.data_ptr()}
    stream = torch.cuda.current_stream().cuda_stream

    for _ in range(args.warmup):
        scene.sun_pass(cam, sun, outs, device=True, stream=stream)
    torch.cuda.synchronize()
    log("timing %d steps" % args.steps)
    trace_ms, shadow_ms, rays, srays = [], [], 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tm = scene.sun_pass(cam, sun, outs, device=True, stream=stream)
        trace_ms.append(tm["trace_ms"])
        shadow_ms.append(tm["shadow_ms"])
        rays += tm["rays_traced"]
        srays += tm["shadow_rays"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = (rays + srays) / elapsed / 1e6
    hit_frac = srays / max(rays, 1)
    avg_trace = float(np.mean(trace_ms))
    avg_shadow = float(np.mean(shadow_ms))
    prim_rays = rays // args.steps
    sh_rays = srays // args.steps

    roof = None
    stats = None
    if os.path.exists(STATS_PATH):
        with open(STATS_PATH) as fh:
            stats = json.load(fh)
        bc, ba = algorithmic_bytes(stats)
        achieved = prim_rays * bc / (avg_trace * 1e-3) / 1e9
        traffic = None
        if os.path.exists(TRAFFIC_PATH):
            with open(TRAFFIC_PATH) as fh:
                traffic = json.load(fh).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                "traffic": traffic, "kernel": "k_primary_sun", "bytes_per_ray": bc,
                "algorithmic_bytes_per_launch": prim_rays * bc, "kernel_ms": avg_trace}
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        try:
            cpu = cpu_baseline(mesh, cam, sun)
        except Exception as e:
            log("cpu baseline failed: %r" % (e,))
    line = {
        "metric": "Mrays/s on 1M-tri mesh (closest hit + sun shadow rays)",
        "value": value,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "icosphere L8 (1310720 tris, R=2km, fbm 20%%), %dx%d primary + 1 sun shadow ray "
                               "per hit, FP64 pose at %.0f km, cached GPU BVH" % (W, H,
                                                                               np.linalg.norm(cam["position"]) / 1e3),
                   "primary_rays": prim_rays, "shadow_rays": sh_rays, "hit_fraction": hit_frac},
        "closest_hit_mrays_s": prim_rays / (avg_trace * 1e-3) / 1e6,
        "any_hit_mrays_s": sh_rays / (avg_shadow * 1e-3) / 1e6 if avg_shadow > 0 else None,
        "kernel_ms": {"k_primary_sun": avg_trace, "k_shadow_sun": avg_shadow},
        "bvh": {"wide_nodes": info["n_wide_nodes"], "leaves": info["n_leaves"], "depth": info["max_depth"],
                "build_ms": info["build_ms"], "sah_cost": info["sah_cost"]},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(json.dumps(line) + "\n")


if __name__ == "__main__":
    main()
