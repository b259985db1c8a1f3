"""Serving: each result is the same alone and in a mixed, batched stream;
admission rejects what can never fit the memory budget."""
import numpy as np
import pytest

from crt import capi, scenes, serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes():
    ms = [scenes.displaced_icosphere(6, seed=1), scenes.geodesic_sphere(40, seed=2)]
    return ms, [capi.Scene(m.vertices, m.indices) for m in ms]


# render launch parameters: defaults, and the graph-captured / lane-camera-ray
# form the batched benchmark picks (three workers capture graphs concurrently)
RENDER_OPTS = [None, dict(samples_per_pass=4, use_graph=True, extend_mode=3)]


@pytest.mark.parametrize("ropts", RENDER_OPTS)
def test_alone_vs_mix_identical(meshes, ropts):
    ms, sc = meshes
    reqs = serving.mixed_workload(ms, 40, rate_hz=1e6, seed=3, max_render=256, max_lidar=128, max_flux=512)
    srv = serving.Server(batching=True, n_workers=3, render_opts=ropts)
    for s in sc:
        srv.add_scene(s)
    mixed, lat, st = serving.run_stream(srv, reqs, realtime=False)
    assert st["completed"] == len(reqs)
    assert st["batches"] < len(reqs)          # requests were batched
    alone_srv = serving.Server(batching=False, render_opts=ropts)
    for s in sc:
        alone_srv.add_scene(s)
    for i in range(0, len(reqs), 3):
        t = alone_srv.submit(reqs[i][1])
        out, _ = alone_srv.result(t)
        for k in out:
            assert np.array_equal(out[k], mixed[i][k]), (i, reqs[i][1]["kind"], k)
    # plain render through the C ABI directly gives the same image too
    for i, (_, r) in enumerate(reqs):
        if r["kind"] == "render":
            rgba, _, _, _ = sc[r["scene"]].render(r["camera"], r["lights"], r["spp"], r["spp"], 0.0, r["bounces"],
                                                  seed=r["seed"])
            if ropts is None:
                assert np.array_equal(rgba, mixed[i]["rgba"])
            else:   # another traversal form: only exact-tie hits may differ
                assert np.mean(np.any(rgba != mixed[i]["rgba"], axis=-1)) < 1e-3
            break
    # and so do plain flash-LiDAR and flux calls (the server stacks single /
    # first / last; its LiDAR scans of one batch share a device buffer and
    # come back through the worker's pinned staging)
    seen = set()
    for i, (_, r) in enumerate(reqs):
        if r["kind"] in seen or r["kind"] == "render":
            continue
        seen.add(r["kind"])
        if r["kind"] == "lidar":
            o = sc[r["scene"]].flash_lidar(r["camera"], r["sub_rays"], r["sigma"], r["seed"])
            assert np.array_equal(np.stack([o["single"], o["first"], o["last"]]), mixed[i]["lidar"])
        else:
            f, t, _, _ = sc[r["scene"]].radiation_pressure(r["sun_dir"], r["pressure"], r["rho_s"], r["rho_d"],
                                                          r["grid"])
            assert np.array_equal(np.concatenate([np.asarray(f), np.asarray(t)]), mixed[i]["force_torque"])
    assert seen == {"lidar", "flux"}
    srv.close()
    alone_srv.close()


def test_admission_rejects_oversized_request(meshes):
    ms, sc = meshes
    srv = serving.Server(batching=True, memory_budget=1 << 20)
    srv.add_scene(sc[0])
    cam = scenes.small_body_camera(ms[0], resolution=512, seed=1)
    with pytest.raises(capi.CrtError) as e:
        srv.submit({"kind": "render", "scene": 0, "camera": cam, "spp": 16, "bounces": 2, "seed": 1, "lights": []})
    assert e.value.code == capi.CRT_ERR_ADMISSION
    # a small request still goes through under the same budget
    small = scenes.small_body_camera(ms[0], resolution=16, seed=1)
    t = srv.submit({"kind": "render", "scene": 0, "camera": small, "spp": 1, "bounces": 1, "seed": 1, "lights": []})
    out, lat = srv.result(t)
    assert out["rgba"].shape == (16, 16, 4) and lat > 0
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_invalid_requests_refused_at_submission(meshes):
    # every option a batch would check is checked at submission: bad requests
    # raise there and never share a batch with valid ones
    ms, sc = meshes
    srv = serving.Server(batching=True, n_workers=2)
    for s in sc:
        srv.add_scene(s)
    cam = scenes.small_body_camera(ms[0], resolution=48, seed=2)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    good = [{"kind": "lidar", "scene": 0, "camera": cam, "sub_rays": 8, "sigma": 5e-4, "seed": 1},
            {"kind": "flux", "scene": 1, "sun_dir": scenes.sun_direction(3), "pressure": 4.56e-6, "rho_s": 0.3,
             "rho_d": 0.6, "grid": 256},
            {"kind": "render", "scene": 0, "camera": cam, "spp": 2, "bounces": 2, "seed": 4, "lights": [sun]}]
    bad = [dict(good[0], sub_rays=64), dict(good[0], sigma=-1.0), dict(good[1], grid=0), dict(good[1], grid=70000),
           dict(good[1], sun_dir=(0.0, 0.0, 0.0)), dict(good[2], spp=0), dict(good[2], scene=7),
           dict(good[2], lights=[{"type": "directional", "intensity": 1.0, "direction": (0.0, 0.0, 0.0)}])]
    tickets = []
    for k in range(3):
        for g in good:
            tickets.append((g, srv.submit(g)))
        for b in bad:
            with pytest.raises((ValueError, TypeError)):
                srv.submit(b)
    for g, tk in tickets:
        out, _ = srv.result(tk)
        if g["kind"] == "lidar":
            o = sc[0].flash_lidar(g["camera"], g["sub_rays"], g["sigma"], g["seed"])
            assert np.array_equal(np.stack([o["single"], o["first"], o["last"]]), out["lidar"])
        elif g["kind"] == "flux":
            f, tq, _, _ = sc[1].radiation_pressure(g["sun_dir"], g["pressure"], g["rho_s"], g["rho_d"], g["grid"])
            assert np.array_equal(np.concatenate([f, tq]), out["force_torque"])
        else:
            rgba, _, _, _ = sc[0].render(g["camera"], g["lights"], 2, 2, 0.0, 2, seed=g["seed"])
            assert np.array_equal(rgba, out["rgba"])
    st = srv.stats()
    assert st["completed"] == len(tickets)
    srv.close()


def test_many_lights_served_and_plain_match_oracle(meshes):
    # no cap on the number of lights (the reference accepts any list):
    # 11 lights in a served request equal the plain render, and the plain
    # render matches the oracle at the same RNG keys
    import oracle
    from render_parity import classify, fp32_eps
    ms, sc = meshes
    rng = np.random.default_rng(4)
    lights = [{"type": "directional", "intensity": 0.1, "direction": scenes.random_unit(rng)} for _ in range(8)] + \
             [{"type": "point", "intensity": 5e8, "position": scenes.random_unit(rng) * 9000.0} for _ in range(3)]
    cam = scenes.small_body_camera(ms[0], resolution=64, seed=8)
    srv = serving.Server(batching=True, n_workers=2)
    srv.add_scene(sc[0])
    req = {"kind": "render", "scene": 0, "camera": cam, "spp": 4, "bounces": 2, "seed": 12, "lights": lights}
    out, _ = srv.result(srv.submit(req))
    srv.close()
    rgba, _, _, _ = sc[0].render(cam, lights, 4, 4, 0.0, 2, seed=12)
    assert np.array_equal(rgba, out["rgba"])
    cpu = oracle.Scene(ms[0].vertices, ms[0].indices)
    info = sc[0].info()
    rgb, _, frag = cpu.render(cam, lights, 4, 4, 0.0, 2, seed=12, offset=info["self_offset"], eps=fp32_eps(info))
    r = classify(rgba[..., :3], rgb, frag)
    print("11 lights: agree %.5f, differ %d, unexplained %d" % (r["agree"], r["differ"], r["unexplained"]))
    assert r["unexplained"] == 0 and r["agree"] > 0.99


def test_tight_budget_counts_retained_workspaces(meshes):
    # a budget of about two renders' working memory, three workers: the
    # workspaces a worker keeps between batches count against the budget,
    # idle workers' are freed when a batch needs the room, and every request
    # completes (no wait on memory that nobody would release)
    ms, sc = meshes
    cam = scenes.small_body_camera(ms[0], resolution=128, seed=1)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    req = {"kind": "render", "scene": 0, "camera": cam, "spp": 4, "bounces": 2, "seed": 1, "lights": [sun]}
    probe = serving.Server(batching=True, n_workers=1)
    probe.add_scene(sc[0])
    probe.result(probe.submit(req))
    probe.close()
    # the estimate of one such request: render_bytes_estimate + image
    px = 128 * 128
    est = px * 4 * (5 * 16 + 8 + 12 + 16 + 48) + px * (16 + 4 + 1 + 4 + 16 + 4) + px * 16
    srv = serving.Server(batching=True, n_workers=3, memory_budget=int(2.2 * est), max_batch=1)
    srv.add_scene(sc[0])
    ts = [srv.submit(dict(req, seed=100 + k)) for k in range(12)]
    for k, t in enumerate(ts):
        out, _ = srv.result(t)
        ref, _, _, _ = sc[0].render(cam, [sun], 4, 4, 0.0, 2, seed=100 + k)
        assert np.array_equal(out["rgba"], ref)
    assert srv.stats()["completed"] == 12
    srv.close()


def test_stats_reset_restarts_the_window(meshes):
    # after a warm-up stream the statistics restart; the server keeps its
    # scenes and workspaces and serves on with the same results
    ms, sc = meshes
    cam = scenes.small_body_camera(ms[0], resolution=64, seed=3)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(4)}
    req = {"kind": "render", "scene": 0, "camera": cam, "spp": 2, "bounces": 2, "seed": 5, "lights": [sun]}
    srv = serving.Server(batching=True, n_workers=2)
    srv.add_scene(sc[0])
    first, _ = srv.result(srv.submit(req))
    srv.result(srv.submit(dict(req, seed=6)))
    st = srv.stats()
    assert st["completed"] == 2 and st["rays"] > 0 and st["wall_s"] > 0
    srv.drain()
    srv.reset_stats()
    st = srv.stats()
    assert st["completed"] == 0 and st["rays"] == 0 and st["batches"] == 0
    assert st["wall_s"] == 0 and st["p50_ms"] == 0
    again, _ = srv.result(srv.submit(req))
    assert np.array_equal(again["rgba"], first["rgba"])
    st = srv.stats()
    assert st["completed"] == 1 and st["batches"] == 1 and st["wall_s"] > 0
    srv.close()


def test_cancel_queued_requests_and_wait_timeout(meshes):
    # one worker, one request at a time: everything behind a large render is
    # still queued right after submission
    ms, sc = meshes
    srv = serving.Server(batching=False)
    srv.add_scene(sc[0])
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    big = {"kind": "render", "scene": 0, "camera": scenes.small_body_camera(ms[0], resolution=1024, seed=3),
           "spp": 16, "bounces": 2, "seed": 5, "lights": [sun]}
    cam = scenes.small_body_camera(ms[0], resolution=48, seed=4)
    small = [{"kind": "render", "scene": 0, "camera": cam, "spp": 2, "bounces": 2, "seed": 10 + k, "lights": [sun]}
             for k in range(8)]
    tb = srv.submit(big)
    ts = [srv.submit(r) for r in small]
    # a zero timeout on a request waiting behind the large one times out, and
    # the ticket stays valid
    with pytest.raises(TimeoutError):
        srv.result(ts[-1], timeout_ms=0.0)
    cancelled = [k for k in range(0, 8, 2) if srv.cancel(ts[k])]
    assert len(cancelled) >= 3          # the large render keeps the worker busy meanwhile
    out_big, _ = srv.result(tb)
    assert out_big["rgba"].shape == (1024, 1024, 4)
    for k, t in enumerate(ts):
        if k in cancelled:
            with pytest.raises(capi.CrtError) as e:
                srv.result(t)
            assert e.value.code == capi.CRT_ERR_CANCELLED
        else:
            out, _ = srv.result(t, timeout_ms=60_000.0)
            ref, _, _, _ = sc[0].render(cam, [sun], 2, 2, 0.0, 2, seed=10 + k)
            assert np.array_equal(out["rgba"], ref)
    st = srv.stats()
    assert st["cancelled"] == len(cancelled) and st["completed"] == 9 - len(cancelled)
    # a finished request's ticket is gone
    with pytest.raises(ValueError):
        srv.cancel(tb)
    srv.close()
