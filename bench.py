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


# Hardware counters of the two kernels at this workload (rocprofv3 --pmc,
# per dispatch): VALU issue, L2 requests and hit rate.
COUNTERS_PATH = os.path.join(ROOT, "profiles", "k_primary_sun_icosphere_L8_2048_counters.json")
HBM_PEAK_GBS = 8000.0


def roofline(stats, prim_rays, sh_rays, t_prim_ms, t_sh_ms, ms_per_step):
    """Roofline of the dominant kernel (k_primary_sun) and of k_shadow_sun.

    `achieved` follows the algorithmic-byte model: reference traversal
    statistics x FP32 node / triangle sizes per ray (see algorithmic_bytes),
    times the rays of one launch, over the kernel's mean launch duration in
    the timed region.  It is a work rate in HBM-equivalent bytes, not
    traffic: the compressed BVH (6.3 MB) and the triangles (63 MB) stay in
    L2 and the 256 MB Infinity Cache, so the measured HBM bytes (`traffic`,
    FETCH_SIZE + WRITE_SIZE scaled by a copy calibration) are a few percent
    of it.  `measured` gives the physical figures from the committed
    counters and names the unit that limits the kernel."""
    bc, ba = algorithmic_bytes(stats)
    traffic = ctr = None
    if os.path.exists(TRAFFIC_PATH):
        with open(TRAFFIC_PATH) as fh:
            traffic = json.load(fh)
    if os.path.exists(COUNTERS_PATH):
        with open(COUNTERS_PATH) as fh:
            ctr = json.load(fh)

    def one(kernel, rays, b, t_ms, hbm):
        ach = rays * b / (t_ms * 1e-3) / 1e9
        out = {"kernel": kernel, "kernel_ms": t_ms, "bytes_per_ray": b, "algorithmic_bytes_per_launch": rays * b,
               "achieved": ach, "frac": ach / HBM_PEAK_GBS, "traffic": hbm}
        m = (ctr or {}).get(kernel)
        if m is not None or hbm is not None:
            meas = {}
            if hbm is not None:
                meas["hbm_GBps"] = hbm / (t_ms * 1e-3) / 1e9
                meas["hbm_frac"] = meas["hbm_GBps"] / HBM_PEAK_GBS
            if m is not None:
                meas["valu_busy_frac"] = m["valu_busy_frac"]
                meas["l2_read_GBps"] = m["l2_read_bytes_per_dispatch"] / (t_ms * 1e-3) / 1e9
                meas["l2_hit_rate"] = m["l2_hit_rate"]
                meas["wave_wait_frac"] = m["wave_wait_frac"]
                meas["limiter"] = m["limiter"]
            out["measured"] = meas
        return out

    prim = one("k_primary_sun", prim_rays, bc, t_prim_ms, (traffic or {}).get("hbm_bytes_per_launch"))
    sh = one("k_shadow_sun", sh_rays, ba, t_sh_ms, ((traffic or {}).get("shadow_kernel") or {}).get(
        "hbm_bytes_per_launch"))
    roof = {"bound": "hbm", "achieved": prim["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": prim["frac"],
            "traffic": prim["traffic"], "kernel": "k_primary_sun", "kernel_ms": t_prim_ms,
            "achieved_is": "algorithmic bytes (reference BVH traversal statistics x FP32 node/triangle bytes per "
                           "ray) per launch / mean launch duration in the timed region: HBM-equivalent work, not "
                           "traffic (the BVH lives in L2 + Infinity Cache)",
            "bytes_per_ray": prim["bytes_per_ray"],
            "algorithmic_bytes_per_launch": prim["algorithmic_bytes_per_launch"],
            "measured": prim.get("measured"), "shadow": sh}
    if prim.get("measured") and "limiter" in prim["measured"]:
        roof["limiter"] = prim["measured"]["limiter"]
    if prim["traffic"] is not None and sh["traffic"] is not None:
        hbm_step = prim["traffic"] + sh["traffic"]
        roof["per_step"] = {"hbm_bytes": hbm_step, "hbm_GBps": hbm_step / (ms_per_step * 1e-3) / 1e9,
                            "hbm_frac": hbm_step / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return roof


def bvh_quality(scenes, scene, mesh, cam, sun, stats):
    """Traversal work per ray of the GPU BVH on the workload's rays, beside
    the reference BVH's statistics for the same rays: every camera ray
    (closest hit), then one sun ray (any hit) per hit from the hit point moved
    max(1e-4, 1e-6 extent) along the inward normal of the hit triangle, as the
    reference's statistics were taken.  The reference's traversal step tests
    two child boxes; here a visit of an 8-wide node tests its valid children."""
    o, d = scenes.camera_rays(cam)
    c = scene.trace_stats(o, d)
    hit = c["tri"] >= 0
    v = np.asarray(mesh.vertices, dtype=np.float64)
    f = np.asarray(mesh.indices, dtype=np.int64)[c["tri"][hit]]
    p0, p1, p2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nin = np.cross(p0 - p1, p2 - p0)
    nin /= np.linalg.norm(nin, axis=1, keepdims=True)
    xyz = o[hit] + c["t"][hit][:, None] * d[hit]
    so = xyz - max(1e-4, 1e-6 * float(np.max(np.abs(v)))) * nin
    sd = np.repeat(np.asarray(sun, dtype=np.float64)[None, :], so.shape[0], axis=0)
    a = scene.trace_stats(so, sd, any_hit=True)
    rc, ra = stats["closest"], stats["shadow"]
    return {"rays": "all %d camera rays (closest) + %d sun rays from their hits (any hit)" % (o.shape[0], so.shape[0]),
            "closest": {"wide_nodes": float(c["nodes"].mean()), "box_tests": float(c["boxes"].mean()),
                        "triangle_tests": float(c["tris"].mean()),
                        "reference_box_tests": 2.0 * rc["traversal_steps_mean"],
                        "reference_triangle_tests": rc["intersections_mean"]},
            "shadow": {"wide_nodes": float(a["nodes"].mean()), "box_tests": float(a["boxes"].mean()),
                       "triangle_tests": float(a["tris"].mean()),
                       "reference_box_tests": 2.0 * ra["traversal_steps_mean"],
                       "reference_triangle_tests": ra["intersections_mean"]}}


def cpu_baseline(mesh, cam, sun, budget_s=12.0, gpu_depth=None, gpu_tri=None):
    """Time a bounded sample of the same workload on the host cores.

    Preferred: the reference tracer itself (oracle/_ref, built from the
    reference sources); otherwise the FP64 oracle restatement.  With the GPU
    pass's depth / triangle outputs (host arrays, [H, W]) it also states the
    GPU result's error against that FP64 CPU tracer on the first sample."""
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
        trace = lambda oo, dd: g.trace(oo, dd)[:2]          # (triangle, t)
    else:
        s = oracle.Scene(mesh.vertices, mesh.indices)
        kind, threads = "port", oracle.num_threads()
        trace = lambda oo, dd: s.trace(oo, dd)[1::-1]       # (triangle, t)
    n = 1 << 16
    total_rays = 0
    t_sum = 0.0
    err = None
    while t_sum < budget_s:
        sel = rng.integers(0, o.shape[0], size=n)
        t0 = time.perf_counter()
        res = trace(o[sel], d[sel])
        t_sum += time.perf_counter() - t0
        if err is None and gpu_depth is not None and gpu_tri is not None:
            err = _error_vs_cpu(res, sel, d, gpu_depth.ravel(), gpu_tri.ravel())
        total_rays += n
        n = min(n * 2, 1 << 20)
    out = {"value": total_rays / t_sum / 1e6, "unit": "Mrays/s", "cores": int(threads), "kind": kind,
           "sample": "%d primary rays drawn uniformly (with replacement) from the 2048^2 camera of the same "
                     "workload, closest hit, cached BVH (build not timed), %.1f s of tracing" % (total_rays, t_sum)}
    if err is not None:
        out["gpu_error_vs_cpu"] = err
    return out


def _error_vs_cpu(res, sel, d, gpu_depth, gpu_tri):
    """GPU closest hits (FP32 traversal, FP64 refinement) against the FP64 CPU
    tracer on the same rays: triangle-id disagreement and the relative
    hit-distance error where both report the same triangle."""
    prim, t = np.asarray(res[0]), np.asarray(res[1])
    cpu_tri = np.where(prim >= 0, prim, -1).astype(np.int64)
    g_tri = gpu_tri[sel].astype(np.int64)
    same = g_tri == cpu_tri
    hit = same & (cpu_tri >= 0)
    dist = t[hit] * np.linalg.norm(d[sel][hit], axis=1)
    rel = np.abs(gpu_depth[sel][hit] - dist) / dist
    return {"rays": int(sel.size), "hits": int(hit.sum()), "id_mismatch_frac": float(1.0 - same.mean()),
            "depth_max_rel_err": float(rel.max()) if rel.size else 0.0,
            "depth_mean_rel_err": float(rel.mean()) if rel.size else 0.0}


def flux_cpu_baseline(mesh, scene, sun, grid=2048, P=4.56e-6, rho_s=0.3, rho_d=0.6):
    """The FP64 oracle's radiation pressure over the same ray grid (its own
    FP64 traversal on the host cores), timed, and the GPU result's force /
    torque error against it (compensated FP64 sums on both sides)."""
    import oracle
    info = scene.info()
    ctr, rad = info["sphere"][:3], info["sphere"][3]
    F, T, nh, _ = scene.radiation_pressure(sun, P, rho_s, rho_d, grid, ctr, rad, (0, 0, 0))
    cs = oracle.Scene(mesh.vertices, mesh.indices)
    t0 = time.perf_counter()
    Fc, Tc, nhc = cs.radiation_pressure(sun, P, rho_s, rho_d, grid, ctr, rad, (0, 0, 0))
    dt = time.perf_counter() - t0
    return {"value": grid * grid / dt / 1e6, "unit": "Mrays/s", "cores": int(oracle.num_threads()), "kind": "port",
            "sample": "the full %d^2 flux grid of the workload (FP64 oracle traversal, BVH build not timed)" % grid,
            "gpu_error_vs_cpu": {"force_rel_err": float(np.linalg.norm(F - Fc) / np.linalg.norm(Fc)),
                                 "torque_rel_err": float(np.linalg.norm(T - Tc) / np.linalg.norm(Tc)),
                                 "hits_gpu": int(nh), "hits_cpu": int(nhc)}}


def run_cornell(args, capi, scenes):
    """The reference's CPU-runnable case: the 36-triangle Cornell box, 256x256
    pinhole camera (60 deg), one 0.25 x 0.25 area light, 16 spp, 4 bounces.
    One step = the plain render() (uint8 + float radiance to host memory) plus
    the intersection / normal / instance passes (host outputs), as a user of
    the drop-in API calls them.  CPU baseline: the reference renderer and
    passes themselves (oracle/_ref) on the host cores, with the GPU passes'
    id agreement and the radiance means beside it."""
    sc = scenes.cornell_box(0)
    v, i, inst = sc.merged()
    scene = capi.Scene(v, i, inst, colors=np.array(sc.colors))
    cam = scenes.cornell_camera()
    light = dict(scenes.cornell_light(), type="area")
    for k in range(args.warmup):
        scene.render(cam, [light], 16, 16, 0.0, 4, seed=k)
        scene.camera_pass(cam)
    rays, rend_ms = 0, []
    t0 = time.perf_counter()
    for k in range(args.steps):
        rgba, u8, _, tm = scene.render(cam, [light], 16, 16, 0.0, 4, seed=1000 + k)
        p = scene.camera_pass(cam)
        rays += tm["rays_traced"] + tm["shadow_rays"] + 256 * 256
        rend_ms.append(tm["total_ms"])
    el = time.perf_counter() - t0
    out = {"metric": "Cornell box render + passes Mrays/s (36 tris, 256x256, 16 spp, 4 bounces)",
           "value": rays / el / 1e6, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "Cornell box (seed 0), 256x256 camera, area light 0.25x0.25, render 16 spp "
                                  "4 bounces (uint8 + radiance to host) + xyz/normal/instance passes (host)",
                      "rays_per_step": rays // args.steps},
           "render_device_ms": float(np.mean(rend_ms))}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cornell_cpu_baseline(sc, cam, light, u8, p)
    return out


def cornell_cpu_baseline(sc, cam, light, gpu_u8, gpu_pass):
    """The reference's own render() and passes (oracle/_ref, OpenMP on the
    host cores) on the same scene: time per step, and the GPU result beside
    it (instance-id agreement of the passes; image mean of the uint8 render,
    which is random on both sides: the reference seeds from random_device)."""
    from oracle import ref as refmod
    if not refmod.available():
        return None
    g = refmod.Group.from_scene(sc)
    runs, t_sum, img, rp = 0, 0.0, None, None
    while t_sum < 10.0 and runs < 5:
        t0 = time.perf_counter()
        img = g.render_u8(cam, [light], 16, 16, 0.0, 4)
        rp = g.passes(cam)
        t_sum += time.perf_counter() - t0
        runs += 1
    gi = gpu_pass["instance"]
    return {"value": runs / t_sum, "unit": "steps/s", "cores": int(refmod.num_threads()), "kind": "reference",
            "sample": "%d full steps (render 256x256 16 spp 4 bounces + passes), %.1f s" % (runs, t_sum),
            "ms_per_step": t_sum * 1e3 / runs,
            "gpu_vs_reference": {"instance_mismatch_frac": float(np.mean(gi != rp["instance"])),
                                 "u8_image_mean_gpu": float(gpu_u8[..., :3].mean()),
                                 "u8_image_mean_reference": float(img[..., :3].mean())}}


def run_batched(args, capi, scenes):
    """256 cameras on the L8 body, sides 2^U(6,10) px, FP64 poses at 20-500 km,
    8 spp, 3 bounces, directional sun: one batched render (autotuned launch
    parameters, hipGraph pass replay) vs per-camera plain launches."""
    from crt import autotune
    mesh = scenes.displaced_icosphere(8, 2000.0, 0.2, 6, 0)
    scene = capi.Scene(mesh.vertices, mesh.indices)
    rng = np.random.default_rng(5)
    cams = []
    for k in range(args.cameras):
        side = int(2 ** rng.uniform(6, 10))
        cams.append(scenes.small_body_camera(mesh, resolution=side, seed=10000 + k, range_km=(20.0, 500.0)))
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    spp, nb = 8, 3
    seeds = list(range(len(cams)))
    if args.render_once:
        # one warm-up and one render of the whole batch on one stream with the
        # default launch parameters (a short, clean run to profile)
        scene.render_batch(cams[:8], [sun], spp, spp, 0.0, nb, seeds=seeds[:8], outputs=False)
        t0 = time.perf_counter()
        _, tmb = scene.render_batch(cams, [sun], spp, spp, 0.0, nb, seeds=seeds, outputs=False)
        t1 = time.perf_counter() - t0
        rays = tmb["rays_traced"] + tmb["shadow_rays"]
        log("batched, one stream, default launch: %.1f ms, %d rays" % (t1 * 1e3, rays))
        return {"metric": "batched render Mrays/s (256 cameras, 8 spp, 3 bounces, L8 body)", "value": rays / t1 / 1e6,
                "unit": "Mrays/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                "config": {"workload": "batched path tracing: %d cameras, spp 8, bounces 3, sun; one render, one "
                                       "stream, default launch parameters" % len(cams)},
                "one_stream_ms": t1 * 1e3, "rays": rays, "timing": tmb}
    log("autotuning on a subset")
    best = autotune.tune(scene, cams, [sun], spp, nb, log=log)
    log("best launch parameters: %s" % best)
    scene.render_batch(cams[:8], [sun], spp, spp, 0.0, nb, seeds=seeds[:8], outputs=False, **best)
    # one batched render of all cameras on one stream
    times1 = []
    for rep in range(3):
        t0 = time.perf_counter()
        _, tmb = scene.render_batch(cams, [sun], spp, spp, 0.0, nb, seeds=seeds, outputs=False, **best)
        times1.append(time.perf_counter() - t0)
    rays = tmb["rays_traced"] + tmb["shadow_rays"]
    t1 = float(np.median(times1))
    log("batched, one stream: median %.1f ms, %d rays" % (t1 * 1e3, rays))
    # the same cameras as three pixel-balanced batches rendered concurrently
    # from three threads on three streams: one batch's bounce and pass
    # boundaries overlap the others' kernels
    import threading
    ng = 3
    px = [int(c["resolution"][0]) * int(c["resolution"][1]) for c in cams]
    groups, load = [[] for _ in range(ng)], [0] * ng
    for k in np.argsort(px)[::-1]:
        g = int(np.argmin(load))
        groups[g].append(int(k))
        load[g] += px[k]
    groups = [sorted(g) for g in groups]
    streams = [capi.Stream() for _ in range(ng)]
    res = [None] * ng

    def work(g):
        res[g] = scene.render_batch([cams[k] for k in groups[g]], [sun], spp, spp, 0.0, nb,
                                    seeds=[seeds[k] for k in groups[g]], outputs=False, stream=streams[g].handle,
                                    **best)[1]

    times = []
    for rep in range(3):
        ths = [threading.Thread(target=work, args=(g,)) for g in range(ng)]
        t0 = time.perf_counter()
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        times.append(time.perf_counter() - t0)
        log("batched run %d (3 streams): %.1f ms" % (rep, times[-1] * 1e3))
    tb = float(np.median(times))
    rays3 = sum(r["rays_traced"] + r["shadow_rays"] for r in res)
    log("batched: median %.1f ms, %d rays" % (tb * 1e3, rays3))
    t0 = time.perf_counter()
    prays = 0
    for k, c in enumerate(cams):
        _, tm = scene.render_batch([c], [sun], spp, spp, 0.0, nb, seeds=[k], outputs=False)
        prays += tm["rays_traced"] + tm["shadow_rays"]
    tp = time.perf_counter() - t0
    log("plain per-camera: %.1f ms" % (tp * 1e3))
    # identity check on a few cameras
    sel = [0, len(cams) // 2, len(cams) - 1]
    outb, _ = scene.render_batch([cams[i] for i in sel], [sun], spp, spp, 0.0, nb, seeds=[seeds[i] for i in sel], **best)
    same = all(np.array_equal(outb[j][0], scene.render_batch([cams[i]], [sun], spp, spp, 0.0, nb, seeds=[seeds[i]])[0][0][0])
               for j, i in enumerate(sel))
    pix = sum(int(c["resolution"][0]) * int(c["resolution"][1]) for c in cams)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: reference renderer on a sample of the cameras; oracle at the same RNG keys")
        cpu = batched_cpu_baseline(scene, mesh, cams, sun, seeds, spp, nb)
    return {"metric": "batched render Mrays/s (256 cameras, 8 spp, 3 bounces, L8 body)", "value": rays3 / tb / 1e6,
            "cpu_baseline": cpu, "one_stream_mrays_s": rays / t1 / 1e6, "one_stream_ms": t1 * 1e3,
            "streams": ng,
            "unit": "Mrays/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "batched path tracing: %d cameras, %d pixels, spp 8, bounces 3, sun; rendered as %d "
                                   "pixel-balanced batches on %d streams at once" % (len(cams), pix, ng, ng)},
            "batched_ms": tb * 1e3, "batched_ms_runs": [t * 1e3 for t in times], "plain_ms": tp * 1e3,
            "rays": rays3,
            "plain_mrays_s": prays / tp / 1e6,
            "speedup_vs_plain": tp / tb, "launch_params": best, "batch_equals_plain": bool(same)}


def batched_cpu_baseline(scene, mesh, cams, sun, seeds, spp, nb, budget_s=10.0):
    """The reference renderer itself (oracle/_ref: do_render's per-pixel loop
    and unidirectional() over the reference BVH, OpenMP on the host cores) on
    the batch's cameras in batch order until ~budget_s, with the sun as a
    PointLight 1e11 m away (the reference has no directional light; the
    direction then differs by < 1e-7 rad); its rays per path are taken from
    the GPU render of the same cameras.  Beside it, the GPU render's
    agreement with the FP64 oracle at identical RNG keys on the batch's
    smallest camera."""
    import oracle
    from oracle import ref as refmod
    out = None
    if refmod.available():
        g = refmod.Group.from_arrays(mesh.vertices, mesh.indices, build_threads=refmod.num_threads())
        pos = (np.asarray(sun["direction"], float) / np.linalg.norm(sun["direction"]) * 1e11).tolist()
        light = {"type": "point", "intensity": 1e22, "position": pos}   # irradiance 1 at the body
        n, paths, t_sum = 0, 0, 0.0
        while t_sum < budget_s and n < len(cams):
            t0 = time.perf_counter()
            g.render_float(cams[n], [light], spp, nb)
            t_sum += time.perf_counter() - t0
            paths += int(cams[n]["resolution"][0]) * int(cams[n]["resolution"][1]) * spp
            n += 1
        _, tm = scene.render_batch(cams[:n], [sun], spp, spp, 0.0, nb, seeds=seeds[:n], outputs=False)
        rpp = (tm["rays_traced"] + tm["shadow_rays"]) / max(paths, 1)
        out = {"value": paths * rpp / t_sum / 1e6, "unit": "Mrays/s", "cores": int(refmod.num_threads()),
               "kind": "reference",
               "sample": "the first %d cameras of the batch (%d paths, %d spp, %d bounces), %.1f s of rendering; "
                         "%.2f rays per path from the GPU render of the same cameras" % (n, paths, spp, nb, t_sum, rpp)}
    k = int(np.argmin([int(c["resolution"][0]) * int(c["resolution"][1]) for c in cams]))
    gpu_img = scene.render_batch([cams[k]], [sun], spp, spp, 0.0, nb, seeds=[seeds[k]])[0][0][0][..., :3]
    cpu_img, _ = oracle.Scene(mesh.vertices, mesh.indices).render(cams[k], [sun], spp, spp, 0.0, nb, seed=seeds[k])
    diff = np.abs(gpu_img - cpu_img)
    close = np.all(diff <= 1e-3 * np.maximum(np.abs(cpu_img), 1e-3) + 1e-5, axis=-1)
    err = {"camera": k, "pixels": int(close.size), "same_key_agreement_frac": float(close.mean()),
           "image_mean_rel_diff": float(abs(gpu_img.mean() - cpu_img.mean()) / max(cpu_img.mean(), 1e-30))}
    if out is None:
        out = {"value": None, "unit": "Mrays/s", "cores": int(oracle.num_threads()), "kind": "port",
               "sample": "reference tracer not built"}
    out["gpu_error_vs_cpu"] = err
    return out


def run_lidar_srp(args, capi, scenes):
    """Same L8 body: 1024^2 flash-LiDAR beams x (1 boresight + 8 Gaussian
    sub-rays, sigma 0.5 mrad) from a seeded FP64 pose, plus a 2048^2 (4.19M
    ray) sun-direction flux grid (rho_s 0.3, rho_d 0.6, compensated FP64
    reduction).  Outputs stay on the device; one step = one scan + one flux
    integration."""
    mesh = scenes.displaced_icosphere(8, 2000.0, 0.2, 6, 0)
    scene = capi.Scene(mesh.vertices, mesh.indices)
    beams = scenes.small_body_camera(mesh, resolution=1024, seed=3)
    sun = scenes.sun_direction(4)
    nb = 1024 * 1024
    bufs = {"single": capi.DeviceBuffer(nb * 8), "first": capi.DeviceBuffer(nb * 8), "last": capi.DeviceBuffer(nb * 8),
            "n_returns": capi.DeviceBuffer(nb)}
    outs = {k: b.ptr for k, b in bufs.items()}
    # the scan and the flux integration of a step run on two streams of the
    # library's runtime and overlap (the flux call returns host results and
    # waits for its own stream only)
    s_lid, s_flux = capi.Stream(), capi.Stream()
    for _ in range(args.warmup):
        scene.flash_lidar(beams, 8, 5e-4, 4, outputs=outs, device=True, stream=s_lid.handle)
        scene.radiation_pressure(sun, 4.56e-6, 0.3, 0.6, 2048, stream=s_flux.handle)
    capi.device_synchronize()
    # per-kernel times (hipEvents on each launch stream), call by call
    lid_ms, flux_ms = [], []
    for _ in range(args.steps):
        tl = scene.flash_lidar(beams, 8, 5e-4, 4, outputs=outs, device=True, stream=s_lid.handle)
        f, tq, nh, tf = scene.radiation_pressure(sun, 4.56e-6, 0.3, 0.6, 2048, stream=s_flux.handle)
        lid_ms.append(tl["trace_ms"])
        flux_ms.append(tf["trace_ms"])
    capi.device_synchronize()
    lrays, frays = tl["rays_traced"] * args.steps, tf["rays_traced"] * args.steps
    # whole-job throughput: K steps, scan enqueued without timing
    t0 = time.perf_counter()
    for _ in range(args.steps):
        scene.flash_lidar(beams, 8, 5e-4, 4, outputs=outs, device=True, stream=s_lid.handle, timing=False)
        f, tq, nh, _ = scene.radiation_pressure(sun, 4.56e-6, 0.3, 0.6, 2048, stream=s_flux.handle)
    capi.device_synchronize()
    el = time.perf_counter() - t0
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: FP64 oracle flux over the same 2048^2 grid")
        cpu = flux_cpu_baseline(mesh, scene, sun)
    return {"metric": "LiDAR + radiation-pressure Mrays/s (1024^2 x 9 flash LiDAR + 2048^2 flux grid, L8 body)",
            "cpu_baseline": cpu,
            "value": (lrays + frays) / el / 1e6, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "flash LiDAR 1024x1024 beams x (1+8) rays sigma 0.5 mrad seed 3/4 + SRP grid "
                                   "2048x2048, rho_s 0.3 rho_d 0.6", "lidar_rays": lrays // args.steps,
                       "flux_rays": frays // args.steps},
            "lidar_mrays_s": lrays / args.steps / (np.mean(lid_ms) * 1e-3) / 1e6,
            "flux_mrays_s": frays / args.steps / (np.mean(flux_ms) * 1e-3) / 1e6,
            "kernel_ms": {"flash_lidar": float(np.mean(lid_ms)), "radiation_pressure": float(np.mean(flux_ms))},
            "force_N": [float(x) for x in f], "torque_Nm": [float(x) for x in tq], "flux_hits": int(nh)}


def run_served(args, capi, scenes):
    """Mixed Poisson stream (60% renders 128^2-1024^2 x 1-16 spp, 30% flash
    LiDAR 64^2-512^2 x 8 sub-rays, 10% flux grids 0.25-4M rays) over 4
    resident geodesic spheres of 1.05M-10M triangles: continuous batching
    server vs plain per-request launches (one at a time), same requests (same
    per-pixel error), saturated throughput and latency at a fixed arrival
    rate."""
    from crt import serving
    ks = [229, 400, 560, 707] if not args.small else [40, 64, 80, 100]
    meshes, sc = [], []
    for i, k in enumerate(ks):
        log("geodesic sphere k=%d" % k)
        m = scenes.geodesic_sphere(k, seed=i)
        meshes.append(m)
        sc.append(capi.Scene(m.vertices, m.indices))
    resident = sum(s.info()["device_bytes"] for s in sc)
    out = {"metric": "served Mrays/s (mixed render/LiDAR/flux stream, 4 resident meshes)", "unit": "Mrays/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "served mixed stream: meshes %s tris, resident %.2f GB" % (
               [m.n_tris for m in meshes], resident / 1e9), "requests": args.requests}}
    # plain = the drop-in API called once per request, one request at a time
    # (the server's batching = 0 mode: one worker, no batching)
    modes = tuple(m for m in (("served", True, args.workers), ("plain", False, 1)) if m[0] in args.modes.split(","))
    # the same render launch parameters for every mode: lane traversal of
    # camera rays; no hipGraph (a request batch renders a few passes, and a
    # capture + instantiate per batch costs more than it saves)
    ropts = dict(extend_mode=3)
    out["render_launch_params"] = ropts
    for mode, batching, nw in modes:
        # steady state, as a long-running server: the server first serves a
        # warm-up stream of the same mix (its workspaces grow to that mix's
        # sizes), then its statistics restart for the timed stream
        srv = serving.Server(batching=batching, n_workers=nw, render_opts=ropts)
        for s in sc:
            srv.add_scene(s)
        warm = serving.mixed_workload(meshes, args.served_warmup, 1e6, seed=99)
        serving.run_stream(srv, warm, realtime=False)
        srv.drain()
        srv.reset_stats()
        reqs = serving.mixed_workload(meshes, args.requests, 1e6, seed=7)
        log("%s: saturated stream of %d requests" % (mode, len(reqs)))
        _, lat, st = serving.run_stream(srv, reqs, realtime=False)
        out[mode] = {"mrays_s": st["mrays_per_s"], "wall_s": st["wall_s"], "workers": nw,
                     "warmup_requests": args.served_warmup, "batches": st["batches"],
                     "rays": st["rays"], "p50_ms_saturated": st["p50_ms"], "p99_ms_saturated": st["p99_ms"]}
        srv.close()
    # Batching without concurrency: the same saturated request set with each
    # (kind, scene) group alone on one worker, groups one after another,
    # largest request first (it runs alone while the rest of its group queues
    # and then goes as one batch), budget and batch limits raised so nothing
    # splits a group further.  Against plain it separates what merged
    # launches gain from what the served mode's concurrent streams gain.
    if args.no_grouped or "plain" not in out or "served" not in out:
        modes = () if args.no_rate else modes
        return _served_rate(args, out, modes, meshes, sc, ropts)
    reqs = serving.mixed_workload(meshes, args.requests, 1e6, seed=7)
    groups = {}
    for _, r in reqs:
        groups.setdefault((r["kind"], r["scene"]), []).append(r)
    size = {"render": lambda r: r["camera"]["resolution"][0] * r["camera"]["resolution"][1] * r["spp"],
            "lidar": lambda r: r["camera"]["resolution"][0] * r["camera"]["resolution"][1],
            "flux": lambda r: r["grid"] ** 2}
    srv = serving.Server(batching=True, n_workers=1, render_opts=ropts, memory_budget=160 << 30, max_batch=1 << 16,
                         max_batch_paths=1 << 40)
    for s in sc:
        srv.add_scene(s)
    grp = {}
    for (kind, sid), rs in sorted(groups.items()):
        rs = sorted(rs, key=size[kind], reverse=True)
        rays0 = srv.stats()["rays"]
        t0 = time.perf_counter()
        ts = [srv.submit(r) for r in rs]
        for t in ts:
            srv.result(t)
        dt = time.perf_counter() - t0
        c = grp.setdefault(kind, {"requests": 0, "rays": 0, "wall_s": 0.0})
        c["requests"] += len(rs)
        c["rays"] += srv.stats()["rays"] - rays0
        c["wall_s"] += dt
    srv.close()
    for c in grp.values():
        c["mrays_s"] = c["rays"] / c["wall_s"] / 1e6
    wall = sum(c["wall_s"] for c in grp.values())
    out["grouped_one_stream"] = {"by_kind": grp, "wall_s": wall,
                                 "mrays_s": sum(c["rays"] for c in grp.values()) / wall / 1e6,
                                 "speedup_vs_plain": out["plain"]["wall_s"] / wall}
    log("grouped, one stream: %.3f s (%s); plain %.3f s, served %.3f s" % (
        wall, ", ".join("%s %.3f s" % (k, c["wall_s"]) for k, c in grp.items()), out["plain"]["wall_s"],
        out["served"]["wall_s"]))
    return _served_rate(args, out, () if args.no_rate else modes, meshes, sc, ropts)


def _served_rate(args, out, modes, meshes, sc, ropts):
    """Latency of each mode at a fixed Poisson arrival rate."""
    from crt import serving
    rate = args.rate
    for mode, batching, nw in modes:
        srv = serving.Server(batching=batching, n_workers=nw, render_opts=ropts)
        for s in sc:
            srv.add_scene(s)
        reqs = serving.mixed_workload(meshes, args.requests, rate, seed=8)
        log("%s: %d requests at %.0f req/s" % (mode, len(reqs), rate))
        _, lat, st = serving.run_stream(srv, reqs, realtime=True)
        out[mode].update({"rate_hz": rate, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                          "mrays_s_at_rate": st["mrays_per_s"]})
        srv.close()
    first = "served" if "served" in out else "plain"
    out["value"] = out[first]["mrays_s"]
    if "served" in out and "plain" in out:
        out["speedup_vs_plain"] = out["served"]["mrays_s"] / max(out["plain"]["mrays_s"], 1e-9)
    return out


# The other workloads of the default run: (key, --workload, extra flags, time
# limit in s).  The Cornell case keeps its CPU baseline (the reference's own
# renderer on the host cores); the others skip theirs to keep the run short.
OTHER_CONFIGS = (
    ("C1_cornell", "cornell", [], 240),
    ("C3_lidar_srp", "lidar_srp", ["--no-cpu-baseline"], 180),
    ("C4_batched", "batched", ["--no-cpu-baseline"], 300),
    ("C5_served", "served", [], 420),
)


def run_other_configs():
    """Each other workload's JSON line, measured by this script in a child
    process of its own (so a failure or a time-out there leaves the headline
    line intact) after the headline's timed region, one after another."""
    import subprocess
    import tempfile
    out = {}
    for key, workload, extra, limit in OTHER_CONFIGS:
        log("%s (--workload %s)" % (key, workload))
        with tempfile.NamedTemporaryFile(suffix=".json", delete=False) as tf:
            path = tf.name
        t0 = time.perf_counter()
        try:
            r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--workload", workload,
                                "--out", path, *extra], stdout=subprocess.DEVNULL, timeout=limit)
            if r.returncode == 0:
                with open(path) as fh:
                    out[key] = json.loads(fh.read())
            else:
                out[key] = {"error": "exit status %d" % r.returncode}
        except subprocess.TimeoutExpired:
            out[key] = {"error": "no result within %d s" % limit}
        except Exception as e:   # unreadable result file
            out[key] = {"error": repr(e)}
        finally:
            if os.path.exists(path):
                os.unlink(path)
        out[key]["run_s"] = time.perf_counter() - t0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--resolution", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    ap.add_argument("--workload", default="c2", choices=["c2", "cornell", "lidar_srp", "batched", "served"],
                    help="c2 (headline, default): closest hit + sun shadow on the 1.31M-tri body; "
                         "cornell: the reference's CPU-runnable Cornell case; lidar_srp: flash LiDAR + flux; "
                         "batched: 256-camera path tracing; served: mixed request stream")
    ap.add_argument("--cameras", type=int, default=256)
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--rate", type=float, default=100.0)
    ap.add_argument("--small", action="store_true", help="served: small meshes (quick check)")
    ap.add_argument("--workers", type=int, default=3, help="served: worker threads (one HIP stream each)")
    ap.add_argument("--served-warmup", type=int, default=300,
                    help="served: requests of the same mix served before the timed stream")
    ap.add_argument("--modes", default="served,plain", help="served: comma list of served, plain")
    ap.add_argument("--no-rate", action="store_true", help="served: skip the fixed-rate latency runs")
    ap.add_argument("--render-once", action="store_true",
                    help="batched: one render of the batch on one stream, default launch parameters (profiling)")
    ap.add_argument("--no-grouped", action="store_true", help="served: skip the one-stream grouped run")
    ap.add_argument("--no-configs", action="store_true",
                    help="c2: skip the other workloads' lines (cornell, lidar_srp, batched, served) that the "
                         "default run adds under other_configs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by image tiles with no exchange); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from crt import capi, scenes

    if args.copy_calibration:
        nbytes = 1 << 30   # 1 GiB read + 1 GiB written per device-to-device copy (a blit kernel)
        x, y = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
        x.fill(1)
        for _ in range(5):
            capi.device_copy(y, x)
        capi.device_synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "copies": 5}))
        return

    if args.workload != "c2":
        run = {"cornell": run_cornell, "lidar_srp": run_lidar_srp, "batched": run_batched,
               "served": run_served}[args.workload]
        line = run(args, capi, scenes)
        print(json.dumps(line), flush=True)
        if args.out:
            with open(args.out, "w") as fh:
                fh.write(json.dumps(line) + "\n")
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
    # device outputs and streams of the HIP runtime the library uses: three
    # streams, each with its own output buffers; consecutive passes rotate
    # over them, so one pass's kernels fill the others' tails (the scene
    # keeps a workspace per stream; 1 / 2 / 3 / 4 streams: 1.14 / 0.79 /
    # 0.77 / 0.77 ms per pass)
    streams = [capi.Stream() for _ in range(3)]
    obufs = [(capi.DeviceBuffer(H * W * 8), capi.DeviceBuffer(H * W * 4), capi.DeviceBuffer(H * W)) for _ in streams]
    outs_s = [{"depth": d.ptr, "tri": t.ptr, "lit": l.ptr} for d, t, l in obufs]
    depth, tri, _ = obufs[0]
    outs, stream = outs_s[0], streams[0].handle

    # warm-up and timed passes have the same structure: passes enqueued back
    # to back, rotating over the streams (device outputs, no per-call
    # synchronisation), so every dispatch of the run overlaps like the timed
    # ones and a profiler's per-kernel average covers the same regime
    ns = len(streams)
    for k in range(max(args.warmup, ns)):
        scene.sun_pass(cam, sun, outs_s[k % ns], device=True, stream=streams[k % ns].handle, timing=False)
    capi.device_synchronize()
    # whole-job throughput: K passes bracketed by device synchronisations of
    # the library's runtime; the library's timing log records a hipEvent pair
    # around each kernel on its launch stream inside this region
    log("timing %d steps" % args.steps)
    capi.timing_log(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        scene.sun_pass(cam, sun, outs_s[k % ns], device=True, stream=streams[k % ns].handle, timing=False)
    capi.device_synchronize()
    elapsed = time.perf_counter() - t0
    capi.timing_log(False)
    n_prim, avg_trace = capi.timing_log_read("k_primary_sun")
    n_sh, avg_shadow = capi.timing_log_read("k_shadow_sun")
    assert n_prim == args.steps and n_sh == args.steps
    # rays of one pass: every pixel casts a primary ray, every hit a shadow ray
    prim_rays = H * W
    sh_rays = int((tri.to_host(np.empty((H, W), np.int32)) >= 0).sum())
    ms_per_step = elapsed * 1e3 / args.steps
    value = (prim_rays + sh_rays) * args.steps / elapsed / 1e6
    hit_frac = sh_rays / max(prim_rays, 1)

    roof = None
    stats = None
    if os.path.exists(STATS_PATH):
        with open(STATS_PATH) as fh:
            stats = json.load(fh)
        roof = roofline(stats, prim_rays, sh_rays, avg_trace, avg_shadow, ms_per_step)
    quality = None
    if stats is not None and args.resolution == 2048:
        log("traversal statistics (untimed)")
        quality = bvh_quality(scenes, scene, mesh, cam, sun, stats)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        try:
            cpu = cpu_baseline(mesh, cam, sun, gpu_depth=depth.to_host(np.empty((H, W))),
                               gpu_tri=tri.to_host(np.empty((H, W), np.int32)))
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
        "kernel_ms": {"k_primary_sun": avg_trace, "k_shadow_sun": avg_shadow,
                      "note": "mean launch durations over the timed region (hipEvents on each launch stream; "
                              "passes on three streams overlap, so kernels share the GPU)"},
        "bvh": {"wide_nodes": info["n_wide_nodes"], "leaves": info["n_leaves"], "depth": info["max_depth"],
                "build_ms": info["build_ms"], "sah_cost": info["sah_cost"], "traversal_per_ray": quality},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if not args.no_configs:
        # the other workloads after the headline's timed region, each in its own process
        del scene, obufs, depth, tri
        line["other_configs"] = run_other_configs()
    print(json.dumps(line), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(json.dumps(line) + "\n")


if __name__ == "__main__":
    main()
