"""GPU parity: path-traced render, LiDAR (batch + flash) and radiation pressure.

* render: same counter-RNG keys as the FP64 oracle -> per-pixel agreement
  except where FP32 vs FP64 geometry sends a path elsewhere (edges,
  terminators); and the reference's own radiance statistics (K independent
  reference renders, tests/golden/cornell_radiance.npz) within Monte-Carlo 3 sigma.
* LiDAR: the reference's batch ranges (tests/golden/lidar_batch.npz) within
  1e-5 relative, apart from edge-grazing beams.
* radiation pressure: FP64 oracle on the same ray grid within 1e-4 relative;
  sphere closed form F = P pi R^2 (1 + 4 rho_d / 9).
"""
import math
import os

import numpy as np
import pytest

import oracle
from crt import capi, scenes
from flash_ref import flash_subrays
from render_parity import classify, fp32_eps

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def cornell():
    sc = scenes.cornell_box(0)
    v, i, inst = sc.merged()
    vn = oracle.vertex_normals(v, i)
    col = np.array(sc.colors)
    gpu = capi.Scene(v, i, inst, vn, colors=col)
    cpu = oracle.Scene(v, i, inst, vn, colors=col)
    return gpu, cpu


@pytest.fixture(scope="module")
def icosphere():
    m = scenes.displaced_icosphere(8)
    return m, capi.Scene(m.vertices, m.indices), oracle.Scene(m.vertices, m.indices)


def _light():
    return dict(scenes.cornell_light(), type="area")


def _same_keys(gpu, cpu, cam, lights, spp, nb, seed, tag):
    # identical RNG keys and the library's self-intersection offset (1e-4,
    # the reference's, at this scale): every pixel that differs must carry a
    # fragile oracle decision (tests/render_parity.py)
    info = gpu.info()
    rgba, u8, spp_g, tm = gpu.render(cam, lights, spp, spp, 0.0, nb, seed=seed)
    rgb_c, spp_c, frag = cpu.render(cam, lights, spp, spp, 0.0, nb, seed=seed, offset=info["self_offset"],
                                    eps=fp32_eps(info))
    r = classify(rgba[..., :3], rgb_c, frag)
    print("%s: agree %.5f, differ %d (closest %d, shadow %d, unexplained %d), image mean rel %.2e (%.2e agreeing)"
          % (tag, r["agree"], r["differ"], r["closest"], r["shadow"], r["unexplained"], r["mean_rel"],
             r["mean_rel_agreeing"]))
    assert r["unexplained"] == 0
    assert r["mean_rel_agreeing"] < 1e-4
    return rgba, u8, spp_g, r


def test_render_matches_oracle_same_keys(cornell):
    gpu, cpu = cornell
    cam = scenes.cornell_camera()
    cam["resolution"] = [128, 128]
    assert gpu.info()["self_offset"] == pytest.approx(1e-4)
    rgba, u8, spp, r = _same_keys(gpu, cpu, cam, [_light()], 16, 4, 7, "Cornell 128^2 x 16 spp x 4 bounces")
    g = rgba[..., :3]
    assert np.all(spp == 16) and np.all(rgba[..., 3] == 1.0)
    assert r["agree"] > 0.99
    # uint8 image = trunc(clamp(256 x, 0, 255)), alpha 255
    assert np.array_equal(u8[..., :3], np.clip(g * 256.0, 0, 255).astype(np.uint8))
    assert np.all(u8[..., 3] == 255)


@pytest.mark.parametrize("res", [(1, 1), (37, 19), (5, 64), (64, 5), (9, 9)])
def test_render_odd_resolutions_match_oracle(cornell, res):
    # images whose sides are not multiples of the 8x8 path tiles (or smaller
    # than one tile), non-square: every pixel is rendered exactly once with
    # its own RNG keys, so same-key renders agree with the FP64 oracle
    gpu, cpu = cornell
    cam = scenes.cornell_camera()
    cam["resolution"] = list(res)
    cam["sensor_size"] = [cam["sensor_size"][0], cam["sensor_size"][0] * res[1] / res[0]]
    rgba, _, spp, _ = _same_keys(gpu, cpu, cam, [_light()], 4, 3, 9, "Cornell %dx%d" % tuple(res))
    assert rgba.shape == (res[1], res[0], 4) and np.all(spp == 4)
    assert np.all(np.isfinite(rgba))
    p = gpu.camera_pass(cam)
    c = cpu.camera_pass(cam)
    mism = p["tri"] != c["tri"]
    if mism.any():
        o, d = oracle.camera_rays(cam)
        sel = mism.ravel()
        assert (cpu.tie_flags(o[sel], d[sel]) | cpu.edge_flags(o[sel], d[sel], p["tri"].ravel()[sel])).all()


def test_render_adaptive_sampling_matches_oracle(cornell):
    gpu, cpu = cornell
    cam = scenes.cornell_camera()
    cam["resolution"] = [64, 64]
    rgba, _, spp, _ = gpu.render(cam, [_light()], 2, 12, 0.05, 3, seed=3)
    rgb_c, spp_c = cpu.render(cam, [_light()], 2, 12, 0.05, 3, seed=3)
    agree = spp == spp_c
    print("adaptive: spp agreement %.4f, mean spp %.2f" % (agree.mean(), spp.mean()))
    assert agree.mean() > 0.95
    assert spp.min() >= 2 and spp.max() <= 12 and spp.min() < 12


def test_render_statistics_match_reference(cornell):
    path = os.path.join(GOLD, "cornell_radiance.npz")
    if not os.path.exists(path):
        pytest.skip("reference radiance statistics not generated")
    g = np.load(path)
    gpu, _ = cornell
    cam = scenes.cornell_camera()
    M = 16
    runs = np.stack([gpu.render(cam, [_light()], 16, 16, 0.0, 4, seed=100 + k)[0][..., :3] for k in range(M)])
    mu_g, var_g = runs.mean(0), runs.var(0, ddof=1)
    K = int(g["k_runs"])
    mu_r, var_r = g["mean"].astype(np.float64), g["var"].astype(np.float64)
    se = np.sqrt(var_g / M + var_r / K)
    ok = se > 0
    z = np.abs(mu_g - mu_r)[ok] / se[ok]
    frac = (z > 3.0).mean()
    zero_both = (~ok) & (np.abs(mu_g - mu_r) > 1e-6)
    print("radiance vs reference: |z|>3 fraction %.4f (MC expectation 0.0027), image means %.5f vs %.5f"
          % (frac, mu_g.mean(), mu_r.mean()))
    assert frac < 0.01
    assert zero_both.sum() == 0
    # whole-image mean within 3 standard errors
    se_img = math.sqrt(var_g.sum() / M + var_r.sum() / K) / mu_g.size
    assert abs(mu_g.mean() - mu_r.mean()) < 3 * se_img + 1e-6


def test_render_point_and_directional_lights(cornell):
    gpu, cpu = cornell
    cam = scenes.cornell_camera()
    cam["resolution"] = [64, 64]
    lights = [{"type": "point", "intensity": 0.05, "position": np.array([0.5, 0.9, 0.5])},
              {"type": "directional", "intensity": 0.2, "direction": np.array([0.3, 0.2, 1.0])}]
    _, _, _, r = _same_keys(gpu, cpu, cam, lights, 4, 2, 1, "Cornell point + directional")
    assert r["agree"] > 0.99


def test_lidar_batch_matches_reference(icosphere):
    m, gpu, cpu = icosphere
    g = np.load(os.path.join(GOLD, "lidar_batch.npz"))
    r = gpu.lidar_batch(g["position"], g["rotation"].reshape(-1, 9))
    ref = g["range"]
    hit_both = (r > 0) & (ref > 0)
    rel = np.abs(r[hit_both] - ref[hit_both]) / ref[hit_both]
    bad = np.where((r > 0) != (ref > 0))[0]
    # a hit-status disagreement is allowed only where the beam's closest-hit
    # decision lies within FP32 resolution of an edge (oracle classification);
    # beam: origin = position, direction R^T (0, 0, -1) = -(third row of R)
    fragile = np.zeros(0, bool)
    if bad.size:
        R = g["rotation"].reshape(-1, 3, 3)[bad]
        from render_parity import fp32_eps
        fragile = cpu.closest_fragile(g["position"][bad], -R[:, 2, :], fp32_eps(gpu.info()))
    print("lidar: %d beams, hit-status disagreements %d (%d near an edge), max rel range err %.2e"
          % (r.size, bad.size, int(fragile.sum()), rel.max()))
    assert int((~fragile).sum()) == 0
    assert rel.max() < 1e-5
    assert np.allclose(r[:8], g["single_range"], rtol=1e-5)


def test_flash_lidar(icosphere):
    m, gpu, cpu = icosphere
    beams = scenes.small_body_camera(m, resolution=128, seed=3)
    out = gpu.flash_lidar(beams, sub_rays=8, sigma=5e-4, seed=4)
    # single return == boresight closest hit (reference LiDAR semantics)
    c = cpu.camera_pass(beams)
    hit = (c["tri"] >= 0) & (out["single"] > 0)
    rel = np.abs(out["single"][hit] - c["depth"][hit]) / c["depth"][hit]
    assert rel.max() < 1e-5
    n = out["n_returns"]
    assert n.max() <= 8
    any_ret = n > 0
    assert np.all(out["first"][any_ret] <= out["last"][any_ret])
    assert np.all(out["first"][~any_ret] == 0) and np.all(out["last"][~any_ret] == 0)
    # sub-ray spread: 0.5 mrad at ~100 km is ~50 m on the surface -> first/last bracket the single return
    both = any_ret & (out["single"] > 0) & (n == 8)
    spread = out["last"][both] - out["first"][both]
    assert np.median(spread) > 0


@pytest.mark.parametrize("K,res", [(1, 48), (8, 48), (23, 47), (63, 40)])
def test_flash_lidar_returns_match_oracle(icosphere, K, res):
    # every sub-ray traced by the FP64 oracle: first / last return = min /
    # max range over the sub-rays that hit, n_returns = their count; beams
    # whose sub-rays graze an edge or vertex (a barycentric within 1e-5 of 0)
    # may legitimately differ and are counted apart.  K sets how many beams
    # share a wave (64 // (K + 1): 32, 7, 2, 1) and res the partial last wave.
    m, gpu, cpu = icosphere
    sigma, seed = 5e-4, 4
    beams = scenes.small_body_camera(m, resolution=res, seed=3)
    out = gpu.flash_lidar(beams, sub_rays=K, sigma=sigma, seed=seed)
    # single return = the boresight ray's closest hit
    c = cpu.camera_pass(beams)
    hit0 = c["tri"] >= 0
    assert ((out["single"] > 0) != hit0).sum() <= 2   # silhouette grazes
    hit0 &= out["single"] > 0
    rel0 = np.abs(out["single"][hit0] - c["depth"][hit0]) / c["depth"][hit0]
    assert rel0.max() < 1e-5
    o, d = flash_subrays(beams, seed, K, sigma)
    t, tri, u, v = cpu.trace(o.reshape(-1, 3), d.reshape(-1, 3))
    t, tri = t.reshape(-1, K), tri.reshape(-1, K)
    w = 1.0 - u - v
    graze = ((np.minimum(np.minimum(u, v), w) < 1e-5) & (tri.ravel() >= 0)).reshape(-1, K).any(axis=1)
    hit = tri >= 0
    rng = np.where(hit, t, np.nan)   # unit directions: range = t
    n_c = hit.sum(axis=1)
    first_c = np.where(n_c > 0, np.nanmin(np.where(hit, rng, np.inf), axis=1), 0.0)
    last_c = np.where(n_c > 0, np.nanmax(np.where(hit, rng, -np.inf), axis=1), 0.0)
    n_g = out["n_returns"].ravel().astype(int)
    fg, lg = out["first"].ravel(), out["last"].ravel()
    ok = ~graze
    print("flash beams %d, with returns %d, grazing %d" % (n_c.size, int((n_c > 0).sum()), int(graze.sum())))
    bad_n = (n_g != n_c) & ok
    assert bad_n.sum() <= max(1, 2e-3 * n_c.size), int(bad_n.sum())
    both = ok & (n_c > 0) & (n_g == n_c)
    assert both.sum() > 0.3 * n_c.size
    assert np.max(np.abs(fg[both] - first_c[both]) / first_c[both]) < 1e-5
    assert np.max(np.abs(lg[both] - last_c[both]) / last_c[both]) < 1e-5
    assert np.all(fg[n_g == 0] == 0) and np.all(lg[n_g == 0] == 0)


def test_flash_lidar_sub_ray_limit(icosphere):
    # a beam's rays share one wavefront: at most 63 sub-rays
    m, gpu, cpu = icosphere
    beams = scenes.small_body_camera(m, resolution=8, seed=3)
    with pytest.raises(ValueError):
        gpu.flash_lidar(beams, sub_rays=64)


def test_radiation_pressure_matches_oracle(icosphere):
    m, gpu, cpu = icosphere
    sun = scenes.sun_direction(2)
    info = gpu.info()
    ctr, rad = info["sphere"][:3], info["sphere"][3]
    P, rs, rd = 4.56e-6, 0.3, 0.6
    F, T, nh, tm = gpu.radiation_pressure(sun, P, rs, rd, 1024, ctr, rad, (0, 0, 0))
    Fc, Tc, nhc = cpu.radiation_pressure(sun, P, rs, rd, 1024, ctr, rad, (0, 0, 0))
    print("SRP: F gpu %s oracle %s; hits %d vs %d" % (F, Fc, nh, nhc))
    assert abs(nh - nhc) <= 2e-4 * nhc
    assert np.linalg.norm(F - Fc) < 1e-4 * np.linalg.norm(Fc)
    print("SRP: force rel err %.2e, torque rel err %.2e" % (np.linalg.norm(F - Fc) / np.linalg.norm(Fc),
                                                            np.linalg.norm(T - Tc) / np.linalg.norm(Tc)))
    assert np.linalg.norm(T - Tc) < 1e-4 * np.linalg.norm(Tc)


def test_radiation_pressure_sphere_closed_form():
    m = scenes.displaced_icosphere(7, radius=1000.0, amplitude=0.0)
    gpu = capi.Scene(m.vertices, m.indices)
    sun = np.array([0.3, -0.5, 0.8])
    sun /= np.linalg.norm(sun)
    P, rs, rd = 4.56e-6, 0.3, 0.6
    F, T, nh, _ = gpu.radiation_pressure(sun, P, rs, rd, 2048, (0, 0, 0), 1010.0, (0, 0, 0))
    Fexp = P * math.pi * 1000.0 ** 2 * (1 + 4 * rd / 9)
    along = -F @ sun
    assert along == pytest.approx(Fexp, rel=2e-3)
    assert np.linalg.norm(F + along * sun) < 1e-3 * Fexp
    # linear in the pressure
    F2, _, _, _ = gpu.radiation_pressure(sun, 2 * P, rs, rd, 2048, (0, 0, 0), 1010.0, (0, 0, 0))
    assert np.allclose(F2, 2 * F, rtol=1e-6)
