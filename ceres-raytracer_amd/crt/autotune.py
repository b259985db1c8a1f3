"""Autotuner for batched rendering launch parameters (crt_batch_opts).

Two stages, timed on a representative subset of the batch (the first
cameras at their own resolution, capped only for very large ones: small
images misjudge the traversal forms):

  1. launch presets: samples per pass, hipGraph replay of a pass, persistent
     traversal kernels (blocks per CU, rays claimed per atomic), the
     traversal form of extension rays and the kernel split;
  2. around the stage-1 winner, one knob at a time: paths per octet in the
     octet bounce kernels (rays_per_octet), live paths per quad in the
     shadow kernel (rays_per_quad) and the camera pixel tile shape
     (tile_width).

The best paths per octet / quad depend on the pass size, which the subset
does not reproduce (256 cameras of the L8 body: the subset picked 4 / 1,
177 ms on the whole batch against 170 ms for the defaults), so when the
subset is not the whole batch the stage-2 winner is confirmed by rendering
the whole batch once with it and once with the stage-1 winner.

Every candidate renders the same images: the RNG is keyed per camera /
pixel / sample, and only the traversal form (lanes vs octets) can flip an
exact tie between two triangles; the stage-2 knobs change no traversal.
So only speed is compared.

Winners are cached per process, keyed by device, library build (a hash of
libcrt_hip.so), scene size, bounces and light count.  A file cache is used
only when asked for (`cache_path`, or the CRT_AUTOTUNE_CACHE environment
variable), so a choice never outlives the build or the process that measured
it unless the caller keeps it on purpose.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from crt import capi

PRESETS = [
    dict(samples_per_pass=0, use_graph=False, persistent=False, extend_mode=1),
    dict(samples_per_pass=0, use_graph=True, persistent=False, extend_mode=1),
    dict(samples_per_pass=0, use_graph=False, persistent=False, extend_mode=3),
    dict(samples_per_pass=0, use_graph=True, persistent=False, extend_mode=2),
    dict(samples_per_pass=0, use_graph=True, persistent=False, extend_mode=3),
    dict(samples_per_pass=2, use_graph=True, persistent=False, extend_mode=1),
    dict(samples_per_pass=2, use_graph=True, persistent=False, extend_mode=3),
    dict(samples_per_pass=4, use_graph=True, persistent=False, extend_mode=3),
    dict(samples_per_pass=4, use_graph=True, persistent=False, extend_mode=2),
    dict(samples_per_pass=4, use_graph=True, persistent=False, extend_mode=3, separate_kernels=2),
    dict(samples_per_pass=4, use_graph=True, persistent=False, extend_mode=3, separate_kernels=1),
    dict(samples_per_pass=0, use_graph=True, persistent=True, blocks_per_cu=32, grab=4, extend_mode=1),
    dict(samples_per_pass=0, use_graph=True, persistent=True, blocks_per_cu=32, grab=4, extend_mode=3),
    dict(samples_per_pass=0, use_graph=True, persistent=True, blocks_per_cu=64, grab=1, extend_mode=1),
]
CANDIDATES = PRESETS   # name kept for callers of the first version

# stage 2: values tried for each knob (0 = the library's choice by pass size)
KNOBS = {
    "rays_per_octet": [0, 1, 2, 4, 8],
    "rays_per_quad": [0, 1, 2, 4],
    "tile_width": [8, 16, 4, 32],
}

_memory = {}
_lib_hash = None


def _device_name():
    """Name of device 0 from the HIP runtime the library uses."""
    import ctypes as C
    try:
        L = capi._hip()
        buf = C.create_string_buffer(256)
        L.hipDeviceGetName.argtypes = [C.c_char_p, C.c_int, C.c_int]
        if L.hipDeviceGetName(buf, 256, 0) == 0:
            return buf.value.decode(errors="replace")
    except Exception:
        pass
    return "unknown"


def _build_id():
    """Hash of the loaded libcrt_hip.so: a rebuilt library is a new key."""
    global _lib_hash
    if _lib_hash is None:
        h = hashlib.sha256()
        with open(capi.LIB_PATH, "rb") as fh:
            for chunk in iter(lambda: fh.read(1 << 20), b""):
                h.update(chunk)
        _lib_hash = h.hexdigest()[:16]
    return _lib_hash


def cache_key(scene: capi.Scene, bounces, n_lights):
    n = scene.n_tris
    bucket = int(np.round(np.log2(max(n, 1))))
    return "v4|%s|%s|tris2^%d|b%d|l%d" % (_device_name(), _build_id(), bucket, bounces, n_lights)


def _cache_path(cache_path):
    return cache_path or os.environ.get("CRT_AUTOTUNE_CACHE") or None


def load_cache(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(path, cache):
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _subset(cams, max_cams, max_res):
    sub = []
    for c in cams[:max_cams]:
        c = dict(c)
        r = min(int(c["resolution"][0]), max_res)
        c["resolution"] = [r, r]
        c["sensor_size"] = [c["sensor_size"][0], c["sensor_size"][1]]
        sub.append(c)
    return sub


def tune(scene: capi.Scene, cams, lights, spp, bounces, max_cams=48, max_res=1024, repeats=1, use_cache=True,
         cache_path=None, stage2=True, confirm=True, log=None):
    """Return the fastest crt_batch_opts (dict) for this kind of batch."""
    key = cache_key(scene, bounces, len(lights))
    path = _cache_path(cache_path)
    if use_cache:
        if key in _memory:
            return dict(_memory[key]["opts"])
        if path:
            disk = load_cache(path)
            if key in disk:
                _memory[key] = disk[key]
                return dict(disk[key]["opts"])
    sub = _subset(cams, max_cams, max_res)
    seeds = list(range(len(sub)))

    def timed(cand, cam_list=sub, n=repeats + 1):
        best = None
        sd = list(range(len(cam_list)))
        for _ in range(n):
            t0 = time.perf_counter()
            scene.render_batch(cam_list, lights, spp, spp, 0.0, bounces, seeds=sd, outputs=False, **cand)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        if log:
            log("autotune %s%s: %.2f ms" % (cand, "" if cam_list is sub else " (whole batch)", best * 1e3))
        return best

    results = [(timed(c), dict(c)) for c in PRESETS]
    results.sort(key=lambda x: x[0])
    best_t, best = results[0]
    stage1 = dict(best)
    if stage2:
        for knob, values in KNOBS.items():
            if knob == "rays_per_octet" and best.get("extend_mode", 0) == 2:
                continue        # no octet bounce kernel in the all-lane form
            if knob == "rays_per_quad" and best.get("persistent", False):
                continue        # the persistent form has its own shadow kernel
            for v in values:
                if best.get(knob, 0 if knob != "tile_width" else 8) == v:
                    continue
                cand = dict(best, **{knob: v})
                t = timed(cand)
                results.append((t, cand))
                if t < best_t:
                    best_t, best = t, cand
        whole = len(sub) == len(cams) and all(
            int(a["resolution"][0]) == int(b["resolution"][0]) and int(a["resolution"][1]) == int(b["resolution"][1])
            for a, b in zip(sub, cams))
        if confirm and best != stage1 and not whole:
            timed(stage1, cams, 1)   # warm-up of the whole batch's workspace
            if timed(stage1, cams, 1) <= timed(best, cams, 1):
                best = stage1
    entry = {"opts": best, "ms": best_t * 1e3, "all": [{"opts": c, "ms": t * 1e3} for t, c in results]}
    if use_cache:
        _memory[key] = entry
        if path:
            disk = load_cache(path)
            disk[key] = entry
            save_cache(path, disk)
    return dict(best)
