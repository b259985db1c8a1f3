"""The batched (C4) and served (C5) workloads against the FP64 oracle.

C4: the first 8 cameras of the batched benchmark's sequence on the L8 body
(sides 2^U(6, 10): 596, 601, 267, 141, 74, 185, 198, 72 px; FP64 poses at
20-500 km), 8 spp, 3 bounces, directional sun, one batched render with
hipGraph pass replay.  Each camera is compared at identical RNG keys with
the oracle (src/crt/do_render.hpp:65-116, src/crt/path_tracing/
unidirectional.hpp:24-96) using the library's self-intersection offset:
>= 99.9 % of pixels within 1e-3 relative, every differing pixel explained
by a fragile oracle decision, and image means over the agreeing pixels
within 1e-4 (no systematic bias).

C5: a mixed request set over two resident geodesic spheres, one of
1 048 820 triangles, with the served config's sizes (renders 128^2-1024^2
at 1-16 spp, flash-LiDAR scans 64^2-512^2 beams, flux grids 0.25M-4M rays),
run batched on three workers: each result equals the same request alone,
and one request of each kind on the 1M-triangle mesh is checked against the
oracle (render at the same keys as above; LiDAR single / first / last
within 1e-5; force and torque each within 1e-4 of their own norm).
"""
import numpy as np
import pytest

import oracle
from crt import capi, scenes, serving
from flash_ref import flash_subrays
from render_parity import classify, fp32_eps

pytestmark = pytest.mark.gpu


def _report(tag, r):
    print("%s: %d px, agree %.5f, differ %d (closest %d, shadow %d, texel %d, unexplained %d), fragile px %.4f, "
          "image mean rel %.2e (%.2e over agreeing pixels)"
          % (tag, r["pixels"], r["agree"], r["differ"], r["closest"], r["shadow"], r["texel"], r["unexplained"],
             r["fragile_pixels"], r["mean_rel"], r["mean_rel_agreeing"]))


def _check_render(r):
    # every differing pixel classified, >= 99.9 % agree, and no systematic
    # bias: the image mean over the agreeing pixels within 1e-4 (a single
    # classified divergent path may move a dim image's whole mean by more)
    assert r["unexplained"] == 0
    assert r["agree"] >= 0.999
    assert r["mean_rel_agreeing"] < 1e-4


@pytest.fixture(scope="module")
def l8():
    m = scenes.displaced_icosphere(8, 2000.0, 0.2, 6, 0)
    return m, capi.Scene(m.vertices, m.indices), oracle.Scene(m.vertices, m.indices)


def test_c4_batched_cameras_match_oracle(l8):
    m, gpu, cpu = l8
    rng = np.random.default_rng(5)
    cams = []
    for k in range(8):
        side = int(2 ** rng.uniform(6, 10))
        cams.append(scenes.small_body_camera(m, resolution=side, seed=10000 + k, range_km=(20.0, 500.0)))
    assert max(c["resolution"][0] for c in cams) >= 512
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    seeds = list(range(8))
    out, tm = gpu.render_batch(cams, [sun], 8, 8, 0.0, 3, seeds=seeds, use_graph=True, samples_per_pass=4)
    info = gpu.info()
    eps = fp32_eps(info)
    tot = {"pixels": 0, "differ": 0, "unexplained": 0}
    for k, c in enumerate(cams):
        rgb, spp, frag = cpu.render(c, [sun], 8, 8, 0.0, 3, seed=seeds[k], offset=info["self_offset"], eps=eps)
        assert np.all(out[k][1] == 8)
        r = classify(out[k][0][..., :3], rgb, frag)
        _report("C4 camera %d (%d^2)" % (k, c["resolution"][0]), r)
        _check_render(r)
        for key in tot:
            tot[key] += r[key]
    print("C4 total: %d px, %d differ (%.2e), %d unexplained" % (tot["pixels"], tot["differ"],
                                                                   tot["differ"] / tot["pixels"], tot["unexplained"]))


@pytest.fixture(scope="module")
def c5():
    ms = [scenes.geodesic_sphere(229, seed=1), scenes.geodesic_sphere(40, seed=2)]
    assert ms[0].n_tris >= 1_000_000
    return ms, [capi.Scene(x.vertices, x.indices) for x in ms]


def _c5_requests(ms):
    rng = np.random.default_rng(11)

    def cam(sid, res, seed):
        return scenes.small_body_camera(ms[sid], resolution=res, seed=seed, range_km=(20.0, 500.0))

    def sun():
        return {"type": "directional", "intensity": 1.0, "direction": scenes.random_unit(rng)}

    reqs = [
        {"kind": "render", "scene": 0, "camera": cam(0, 1024, 21), "spp": 1, "bounces": 2, "seed": 21,
         "lights": [sun()]},
        {"kind": "render", "scene": 0, "camera": cam(0, 256, 22), "spp": 16, "bounces": 2, "seed": 22,
         "lights": [sun()]},
        {"kind": "render", "scene": 1, "camera": cam(1, 128, 23), "spp": 4, "bounces": 2, "seed": 23,
         "lights": [sun()]},
        {"kind": "lidar", "scene": 0, "camera": cam(0, 512, 24), "sub_rays": 8, "sigma": 5e-4, "seed": 24},
        {"kind": "lidar", "scene": 1, "camera": cam(1, 64, 25), "sub_rays": 8, "sigma": 5e-4, "seed": 25},
        {"kind": "flux", "scene": 0, "sun_dir": scenes.random_unit(rng), "pressure": 4.56e-6, "rho_s": 0.3,
         "rho_d": 0.6, "grid": 2048},
        {"kind": "flux", "scene": 0, "sun_dir": scenes.random_unit(rng), "pressure": 4.56e-6, "rho_s": 0.3,
         "rho_d": 0.6, "grid": 512},
        {"kind": "flux", "scene": 1, "sun_dir": scenes.random_unit(rng), "pressure": 4.56e-6, "rho_s": 0.3,
         "rho_d": 0.6, "grid": 700},
        {"kind": "render", "scene": 0, "camera": cam(0, 384, 26), "spp": 8, "bounces": 2, "seed": 26,
         "lights": [sun()]},
        {"kind": "lidar", "scene": 0, "camera": cam(0, 200, 27), "sub_rays": 8, "sigma": 5e-4, "seed": 27},
    ]
    return reqs


def test_c5_served_requests_match_alone_and_oracle(c5):
    ms, sc = c5
    reqs = _c5_requests(ms)
    srv = serving.Server(batching=True, n_workers=3)
    for s in sc:
        srv.add_scene(s)
    mixed, lat, st = serving.run_stream(srv, [(0.0, r) for r in reqs], realtime=False)
    assert st["completed"] == len(reqs)
    assert st["batches"] < len(reqs)
    alone = serving.Server(batching=False)
    for s in sc:
        alone.add_scene(s)
    for i, r in enumerate(reqs):
        out, _ = alone.result(alone.submit(r))
        for k in out:
            assert np.array_equal(out[k], mixed[i][k]), (i, r["kind"], k)
    srv.close()
    alone.close()
    print("C5 mix: %d requests in %d batches, p50 %.1f ms, p99 %.1f ms"
          % (len(reqs), st["batches"], st["p50_ms"], st["p99_ms"]))

    cpu = oracle.Scene(ms[0].vertices, ms[0].indices)
    info = sc[0].info()
    eps = fp32_eps(info)
    # renders on the 1M-triangle mesh at identical RNG keys
    for i, r in enumerate(reqs):
        if r["kind"] != "render" or r["scene"] != 0:
            continue
        rgb, _, frag = cpu.render(r["camera"], r["lights"], r["spp"], r["spp"], 0.0, r["bounces"], seed=r["seed"],
                                  offset=info["self_offset"], eps=eps)
        res = classify(mixed[i]["rgba"][..., :3], rgb, frag)
        _report("C5 render %d (%d^2, %d spp)" % (i, r["camera"]["resolution"][0], r["spp"]), res)
        _check_render(res)
    # flash LiDAR: every sub-ray traced by the oracle
    for i, r in enumerate(reqs):
        if r["kind"] != "lidar" or r["scene"] != 0:
            continue
        K = r["sub_rays"]
        single, first, last = mixed[i]["lidar"]
        c = cpu.camera_pass(r["camera"])
        hit0 = (c["tri"] >= 0) & (single > 0)
        assert ((single > 0) != (c["tri"] >= 0)).sum() <= 2
        assert np.max(np.abs(single[hit0] - c["depth"][hit0]) / c["depth"][hit0]) < 1e-5
        o, d = flash_subrays(r["camera"], r["seed"], K, r["sigma"])
        t, tri, u, v = cpu.trace(o.reshape(-1, 3), d.reshape(-1, 3))
        hit = (tri >= 0).reshape(-1, K)
        t = t.reshape(-1, K)
        w = 1.0 - u - v
        graze = ((np.minimum(np.minimum(u, v), w) < 1e-5) & (tri >= 0)).reshape(-1, K).any(axis=1)
        n_c = hit.sum(axis=1)
        first_c = np.where(n_c > 0, np.min(np.where(hit, t, np.inf), axis=1), 0.0)
        last_c = np.where(n_c > 0, np.max(np.where(hit, t, -np.inf), axis=1), 0.0)
        fg, lg = first.ravel(), last.ravel()
        n_g = ((fg > 0).astype(int))
        ok = ~graze & (n_c > 0) & (n_g > 0)
        assert ((n_c > 0) != (n_g > 0))[~graze].sum() <= 2
        assert np.max(np.abs(fg[ok] - first_c[ok]) / first_c[ok]) < 1e-5
        assert np.max(np.abs(lg[ok] - last_c[ok]) / last_c[ok]) < 1e-5
        print("C5 lidar %d: %d beams, %d with returns, %d grazing" % (i, n_c.size, int((n_c > 0).sum()),
                                                                     int(graze.sum())))
    # flux grids: force and torque each within 1e-4 of their own norm
    for i, r in enumerate(reqs):
        if r["kind"] != "flux" or r["scene"] != 0:
            continue
        Fc, Tc, _ = cpu.radiation_pressure(r["sun_dir"], r["pressure"], r["rho_s"], r["rho_d"], r["grid"],
                                           info["sphere"][:3], info["sphere"][3], (0.0, 0.0, 0.0))
        F, T = mixed[i]["force_torque"][:3], mixed[i]["force_torque"][3:]
        ef = np.linalg.norm(F - Fc) / np.linalg.norm(Fc)
        et = np.linalg.norm(T - Tc) / np.linalg.norm(Tc)
        print("C5 flux %d (grid %d): force rel err %.2e, torque rel err %.2e" % (i, r["grid"], ef, et))
        assert ef < 1e-4 and et < 1e-4
