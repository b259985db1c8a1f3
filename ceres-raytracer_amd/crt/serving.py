"""Serving: resident scenes, memory-based admission, continuous batching.

Thin Python view of the crt_server_* C ABI plus the mixed-request workload of
the served benchmark (Poisson arrivals of renders, LiDAR scans and flux
grids).  The server itself (admission, batching, worker streams) is native.
"""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from crt import capi


class Server:
    def __init__(self, batching=True, n_workers=3, memory_budget=16 << 30, max_batch=64,
                 max_batch_paths=1 << 23, render_opts=None):
        """render_opts: crt_batch_opts fields (dict, e.g. from crt.autotune)
        used for every render launch, batched or plain."""
        capi.init()
        ro = dict(render_opts or {})
        bo = capi.BatchOpts(int(ro.get("samples_per_pass", 0)), int(bool(ro.get("use_graph", False))),
                            int(bool(ro.get("persistent", False))), int(ro.get("blocks_per_cu", 0)),
                            int(ro.get("grab", 0)), int(ro.get("extend_mode", 0)),
                            int(ro.get("separate_kernels", 0)), int(ro.get("coherence_sort", 0)),
                            int(ro.get("rays_per_octet", 0)), int(ro.get("rays_per_quad", 0)),
                            int(ro.get("tile_width", 0)), int(ro.get("fixed_share", 0)))
        o = capi.ServerOpts(int(memory_budget), int(n_workers), int(max_batch), int(max_batch_paths),
                            int(bool(batching)), 0, bo)
        h = C.c_void_p()
        capi._check(capi.load().crt_server_create(C.byref(o), C.byref(h)))
        self.h = h
        self._keep = {}

    def add_scene(self, scene: capi.Scene) -> int:
        sid = C.c_uint32()
        capi._check(capi.load().crt_server_add_scene(self.h, scene.h, C.byref(sid)))
        self._keep[("scene", sid.value)] = scene
        return sid.value

    def submit(self, req: dict) -> int:
        """req: kind ('render'|'lidar'|'flux'), scene, and the kind's fields.
        Output arrays are allocated here and returned by result()."""
        r = capi.Request()
        kind = req["kind"]
        r.kind = {"render": capi.REQ_RENDER, "lidar": capi.REQ_LIDAR, "flux": capi.REQ_FLUX}[kind]
        r.scene = int(req["scene"])
        out = {}
        if kind in ("render", "lidar"):
            r.camera = capi.make_camera(req["camera"])
            h, w = int(np.floor(req["camera"]["resolution"][1])), int(np.floor(req["camera"]["resolution"][0]))
        if kind == "render":
            r.spp = int(req["spp"])
            r.bounces = int(req.get("bounces", 1))
            r.seed = int(req.get("seed", 0))
            lights = req.get("lights", [])
            r.n_lights = len(lights)
            if len(lights) <= 2:
                for k, l in enumerate(lights):
                    r.lights[k] = capi.make_light(l)
            else:   # any number of lights through light_list (copied at submission)
                arr = (capi.Light * len(lights))(*[capi.make_light(l) for l in lights])
                r.light_list = C.cast(arr, C.c_void_p)
            out["rgba"] = np.zeros((h, w, 4), dtype=np.float32)
            r.rgba = out["rgba"].ctypes.data
        elif kind == "lidar":
            r.sub_rays = int(req.get("sub_rays", 8))
            r.sigma = float(req.get("sigma", 5e-4))
            r.seed = int(req.get("seed", 0))
            out["lidar"] = np.zeros((3, h, w))
            r.lidar = out["lidar"].ctypes.data
        else:
            f = capi.FluxOpts()
            f.sun_dir[:] = [float(x) for x in req["sun_dir"]]
            f.pressure, f.rho_s, f.rho_d = float(req["pressure"]), float(req["rho_s"]), float(req["rho_d"])
            f.grid = int(req["grid"])
            f.radius = 0.0
            f.ref_point[:] = [float(x) for x in req.get("ref_point", (0.0, 0.0, 0.0))]
            r.flux = f
            out["force_torque"] = np.zeros(6)
            r.force_torque = out["force_torque"].ctypes.data
        t = C.c_uint64()
        capi._check(capi.load().crt_server_submit(self.h, C.byref(r), C.byref(t)))
        self._keep[t.value] = out
        return t.value

    def result(self, ticket: int, timeout_ms=None):
        """(outputs, latency ms) of a request.  With timeout_ms, raises
        TimeoutError if it has not finished by then (ask again later); a
        cancelled request raises CrtError (CRT_ERR_CANCELLED)."""
        lat = C.c_double()
        t = float("inf") if timeout_ms is None else float(timeout_ms)
        code = capi.load().crt_server_wait_for(self.h, ticket, t, C.byref(lat))
        if code == capi.CRT_ERR_TIMEOUT:
            capi._check(code)
        out = self._keep.pop(ticket, None)
        capi._check(code)
        return out, lat.value

    def cancel(self, ticket: int) -> bool:
        """Cancel a request that is still queued; False if it is already
        running or finished (its result then comes as usual)."""
        code = capi.load().crt_server_cancel(self.h, ticket)
        if code == capi.CRT_ERR_STATE:
            return False
        capi._check(code)
        return True

    def drain(self):
        capi._check(capi.load().crt_server_drain(self.h))

    def reset_stats(self):
        """Start the statistics afresh (e.g. after a warm-up stream); only
        while nothing is queued or running."""
        capi._check(capi.load().crt_server_stats_reset(self.h))

    def stats(self) -> dict:
        s = capi.ServerStats()
        capi._check(capi.load().crt_server_stats_get(self.h, C.byref(s)))
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            capi.load().crt_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mixed_workload(meshes, n_requests, rate_hz, seed=0, max_render=1024, max_lidar=512, max_flux=2048):
    """Poisson stream: 60% renders (128^2-1024^2, 1-16 spp), 30% flash-LiDAR
    scans (64^2-512^2 beams, 8 sub-rays), 10% flux grids (0.25-4M rays), over
    the given resident meshes; poses at 20-500 km."""
    from crt import scenes
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        sid = int(rng.integers(0, len(meshes)))
        u = rng.uniform()
        cam_seed = int(rng.integers(0, 2 ** 31))
        sun = scenes.random_unit(rng)
        if u < 0.6:
            res = int(2 ** rng.uniform(7, np.log2(max_render)))
            cam = scenes.small_body_camera(meshes[sid], resolution=res, seed=cam_seed, range_km=(20.0, 500.0))
            req = {"kind": "render", "scene": sid, "camera": cam, "spp": int(rng.integers(1, 17)), "bounces": 2,
                   "seed": cam_seed, "lights": [{"type": "directional", "intensity": 1.0, "direction": sun}]}
        elif u < 0.9:
            res = int(2 ** rng.uniform(6, np.log2(max_lidar)))
            cam = scenes.small_body_camera(meshes[sid], resolution=res, seed=cam_seed, range_km=(20.0, 500.0))
            req = {"kind": "lidar", "scene": sid, "camera": cam, "sub_rays": 8, "sigma": 5e-4, "seed": cam_seed}
        else:
            grid = int(2 ** rng.uniform(9, np.log2(max_flux)))
            req = {"kind": "flux", "scene": sid, "sun_dir": sun, "pressure": 4.56e-6, "rho_s": 0.3, "rho_d": 0.6,
                   "grid": grid}
        reqs.append((t, req))
    return reqs


def run_stream(server: Server, reqs, realtime=True):
    """Submit (arrival_time, request) pairs at their arrival times from a
    feeder thread; returns (results by index, latencies ms, stats)."""
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def feed():
        for i, (ta, req) in enumerate(reqs):
            if realtime:
                dt = ta - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
            tickets[i] = server.submit(req)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    results, lat = [], []
    for t in tickets:
        out, l = server.result(t)
        results.append(out)
        lat.append(l)
    return results, np.array(lat), server.stats()
