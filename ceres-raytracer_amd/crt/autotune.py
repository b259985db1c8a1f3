"""Autotuner for batched rendering launch parameters.

Tunes crt_batch_opts -- samples per pass, hipGraph replay of a pass,
persistent traversal kernels, persistent blocks per CU, rays claimed per
atomic and the traversal form of extension rays -- by timing a representative
subset of the batch (the first cameras at their own resolution, capped only
for very large ones: small images misjudge the traversal forms) and keeps the
winner in a JSON cache keyed by device, scene size, bounces and light count.
Every candidate renders the same images (the RNG is keyed per camera/pixel/
sample; only the traversal form can flip an exact tie between two
triangles), so only speed is compared.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from crt import capi

CACHE = os.environ.get("CRT_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "crt", "autotune.json"))

CANDIDATES = [
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


def _key(scene: capi.Scene, bounces, n_lights):
    n = scene.n_tris
    bucket = int(np.round(np.log2(max(n, 1))))
    return "v3|%s|tris2^%d|b%d|l%d" % (_device_name(), bucket, bounces, n_lights)


def load_cache():
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(cache):
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def tune(scene: capi.Scene, cams, lights, spp, bounces, max_cams=48, max_res=1024, repeats=1, use_cache=True,
         log=None):
    """Return the fastest crt_batch_opts (dict) for this kind of batch."""
    key = _key(scene, bounces, len(lights))
    cache = load_cache() if use_cache else {}
    if key in cache:
        return cache[key]["opts"]
    sub = []
    for c in cams[:max_cams]:
        c = dict(c)
        r = min(int(c["resolution"][0]), max_res)
        scale = r / float(c["resolution"][0])
        c["resolution"] = [r, r]
        c["sensor_size"] = [c["sensor_size"][0], c["sensor_size"][1]]
        sub.append(c)
    results = []
    for cand in CANDIDATES:
        best = None
        for _ in range(repeats + 1):
            t0 = time.perf_counter()
            _, tm = scene.render_batch(sub, lights, spp, spp, 0.0, bounces, seeds=list(range(len(sub))),
                                       outputs=False, **cand)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        results.append((best, cand))
        if log:
            log("autotune %s: %.2f ms" % (cand, best * 1e3))
    results.sort(key=lambda x: x[0])
    winner = results[0][1]
    cache[key] = {"opts": winner, "ms": results[0][0] * 1e3,
                  "all": [{"opts": c, "ms": t * 1e3} for t, c in results]}
    if use_cache:
        save_cache(cache)
    return winner
