"""ctypes binding of oracle/_ref/libcrtref.so -- the reference CPU tracer.

TEST INFRASTRUCTURE ONLY.  The library is built by oracle/Makefile.ref from
the reference's own header sources (only where they exist); the .so then
travels with the tree so the CPU baseline can time the reference itself.
Meshes are handed to the reference as OBJ files written with %.9g, which its
strtof-based loader reads back bit-exactly (lib/model_loaders/obj.hpp:75-77).
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_ref", "libcrtref.so")
_lib = None
_p = C.c_void_p


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("reference harness not built (oracle/Makefile.ref needs the reference sources)")
        L = C.CDLL(LIB)
        L.ref_quiet.argtypes = [C.c_int]
        L.ref_num_threads.restype = C.c_int
        L.ref_group_create.restype = _p
        L.ref_group_create.argtypes = [C.c_int, C.POINTER(C.c_char_p), _p, _p, _p, _p, _p, C.c_int]
        L.ref_group_destroy.argtypes = [_p]
        L.ref_group_ntris.restype = C.c_int64
        L.ref_group_ntris.argtypes = [_p]
        L.ref_group_tris.argtypes = [_p, _p, _p, _p]
        L.ref_group_bvh_info.argtypes = [_p, _p, _p]
        L.ref_trace.argtypes = [_p, C.c_int64, _p, _p, _p, _p, C.c_int, _p, _p, _p, _p, _p, _p]
        L.ref_pass.argtypes = [_p, _p, _p, _p, _p]
        L.ref_render_u8.argtypes = [_p, _p, C.c_int, _p, C.c_int, C.c_int, C.c_double, C.c_int, _p]
        L.ref_render_float.argtypes = [_p, _p, C.c_int, _p, C.c_int, C.c_int, C.c_int, _p]
        L.ref_lidar.restype = C.c_double
        L.ref_lidar.argtypes = [_p, _p, _p, C.c_int]
        L.ref_batch_lidar.argtypes = [_p, C.c_int, _p, _p, C.c_int, _p]
        L.ref_quiet(1)
        _lib = L
    return _lib


def num_threads() -> int:
    return lib().ref_num_threads()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def cam_params(cam: dict) -> np.ndarray:
    return np.array(list(np.asarray(cam["position"], float).reshape(3)) +
                    list(np.asarray(cam["rotation"], float).reshape(9)) +
                    [float(cam["focal_length"])] + [float(x) for x in cam["resolution"]] +
                    [float(x) for x in cam["sensor_size"]] + [1.0 if cam.get("z_positive") else 0.0])


def light_params(lights) -> np.ndarray:
    rows = []
    for l in lights:
        t = {"point": 0.0, "area": 1.0}[l.get("type", "area")]
        rows.append([t, float(l["intensity"])] + list(np.asarray(l.get("position", np.zeros(3)), float)) +
                    list(np.asarray(l.get("rotation", np.eye(3)), float).reshape(9)) +
                    [float(x) for x in l.get("size", [0.0, 0.0])])
    return np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(-1, 16))


class Group:
    """Reference BodyFixedGroup (SweepSAH + PRO + NLO BVH built on the CPU)."""

    def __init__(self, obj_paths, positions=None, rotations=None, scales=None, colors=None, smooth=None,
                 build_threads=1):
        n = len(obj_paths)
        self.paths = [os.fsencode(p) for p in obj_paths]
        arr = (C.c_char_p * n)(*self.paths)
        pos = np.ascontiguousarray(np.zeros((n, 3)) if positions is None else positions, dtype=np.float64)
        rot = np.ascontiguousarray(np.tile(np.eye(3), (n, 1, 1)) if rotations is None else rotations,
                                   dtype=np.float64)
        sc = np.ascontiguousarray(np.ones(n) if scales is None else scales, dtype=np.float64)
        col = np.ascontiguousarray(np.ones((n, 3)) if colors is None else colors, dtype=np.float32)
        sm = np.ascontiguousarray(np.zeros(n) if smooth is None else smooth, dtype=np.int32)
        self.h = lib().ref_group_create(n, arr, _ptr(pos), _ptr(rot), _ptr(sc), _ptr(col), _ptr(sm), build_threads)

    @classmethod
    def from_arrays(cls, vertices, indices, build_threads=1):
        d = tempfile.mkdtemp(prefix="crtref_")
        path = os.path.join(d, "mesh.obj")
        sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "ceres-raytracer_amd"))
        from crt.scenes import Mesh, write_obj
        write_obj(path, Mesh(np.asarray(vertices, np.float32), np.asarray(indices, np.uint32)))
        g = cls([path], build_threads=build_threads)
        os.remove(path)
        os.rmdir(d)
        return g

    @classmethod
    def from_scene(cls, scene, build_threads=1):
        """scene: crt.scenes.Scene (one OBJ per entity, ids 1..n in order)."""
        d = tempfile.mkdtemp(prefix="crtref_")
        sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "ceres-raytracer_amd"))
        from crt.scenes import write_obj
        paths = []
        for k, m in enumerate(scene.meshes):
            p = os.path.join(d, "e%d.obj" % k)
            write_obj(p, m)
            paths.append(p)
        g = cls(paths, colors=np.array(scene.colors), smooth=np.array(scene.smooth, dtype=np.int32),
                build_threads=build_threads)
        for p in paths:
            os.remove(p)
        os.rmdir(d)
        return g

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_group_destroy(self.h)
            self.h = None

    @property
    def n_tris(self):
        return lib().ref_group_ntris(self.h)

    def tris(self):
        n = self.n_tris
        p = np.zeros((n, 9))
        inst = np.zeros(n, dtype=np.uint32)
        vn = np.zeros((n, 9))
        lib().ref_group_tris(self.h, _ptr(p), _ptr(inst), _ptr(vn))
        return p, inst, vn

    def bvh_info(self):
        nc = C.c_int64(0)
        sah = C.c_double(0)
        lib().ref_group_bvh_info(self.h, C.byref(nc), C.byref(sah))
        return {"node_count": nc.value, "sah_cost": sah.value}

    def trace(self, o, d, tmin=None, tmax=None, any_hit=False, stats=False):
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        n = o.shape[0]
        prim = np.empty(n, dtype=np.int64)
        t = np.empty(n)
        u = np.empty(n)
        v = np.empty(n)
        st = np.empty(n, dtype=np.uint32) if stats else None
        it = np.empty(n, dtype=np.uint32) if stats else None
        tmin = None if tmin is None else np.ascontiguousarray(tmin, dtype=np.float64)
        tmax = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float64)
        lib().ref_trace(self.h, n, _ptr(o), _ptr(d), _ptr(tmin), _ptr(tmax), int(any_hit), _ptr(prim), _ptr(t),
                        _ptr(u), _ptr(v), _ptr(st), _ptr(it))
        if stats:
            return prim, t, u, v, st, it
        return prim, t, u, v

    def passes(self, cam):
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        xyz = np.zeros((h, w, 3))
        inst = np.zeros((h, w), dtype=np.uint32)
        nrm = np.zeros((h, w, 3))
        lib().ref_pass(self.h, _ptr(cam_params(cam)), _ptr(xyz), _ptr(inst), _ptr(nrm))
        return {"xyz": xyz, "instance": inst, "normal": nrm}

    def render_u8(self, cam, lights, min_samples, max_samples, noise_threshold, num_bounces):
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        out = np.zeros((h, w, 4), dtype=np.uint8)
        lp = light_params(lights)
        lib().ref_render_u8(self.h, _ptr(cam_params(cam)), lp.shape[0], _ptr(lp), min_samples, max_samples,
                            noise_threshold, num_bounces, _ptr(out))
        return out

    def render_float(self, cam, lights, spp, num_bounces, threads=0):
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        out = np.zeros((h, w, 3), dtype=np.float32)
        lp = light_params(lights)
        lib().ref_render_float(self.h, _ptr(cam_params(cam)), lp.shape[0], _ptr(lp), spp, num_bounces, threads,
                               _ptr(out))
        return out

    def lidar(self, pos, rot, z_positive=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(9)
        return lib().ref_lidar(self.h, _ptr(pos), _ptr(rot), int(z_positive))

    def batch_lidar(self, pos, rot, z_positive=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
        out = np.zeros(pos.shape[0])
        lib().ref_batch_lidar(self.h, pos.shape[0], _ptr(pos), _ptr(rot), int(z_positive), _ptr(out))
        return out
