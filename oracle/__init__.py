"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP core.  Only tests/, the
smoke check in __graft_entry__ and bench.py's cpu_baseline leg may import it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


class Camera(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("rotation", C.c_double * 9), ("focal_length", C.c_double),
                ("resolution", C.c_double * 2), ("sensor_size", C.c_double * 2), ("z_positive", C.c_int32),
                ("pad_", C.c_int32)]


class Material(C.Structure):
    # src/crt/materials/material.hpp:90-184; type 0 ColoredLambertian,
    # 1 TexturedLambertian, 2 TexturedBlinnPhong, 3 Mirror
    _fields_ = [("type", C.c_uint32), ("texture", C.c_int32), ("spec_texture", C.c_int32), ("alpha", C.c_float),
                ("color", C.c_float * 3), ("coeffs", C.c_float * 3)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p)]


MATERIAL_TYPES = {"lambertian": 0, "textured_lambertian": 1, "blinn_phong": 2, "mirror": 3}


def make_material(m: dict) -> Material:
    x = Material()
    x.type = MATERIAL_TYPES[m.get("type", "lambertian")]
    x.texture = int(m.get("texture", -1))
    x.spec_texture = int(m.get("spec_texture", -1))
    x.alpha = float(m.get("alpha", 24.0))
    x.color[:] = [float(c) for c in m.get("color", (1.0, 1.0, 1.0))]
    x.coeffs[:] = [float(c) for c in m.get("coeffs", (0.0, 1.0, 0.0))]
    return x


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad_", C.c_int32), ("intensity", C.c_double), ("position", C.c_double * 3),
                ("rotation", C.c_double * 9), ("size", C.c_double * 2), ("direction", C.c_double * 3)]


_p = C.c_void_p


def _declare(L):
    L.crt_oracle_scene_create.restype = _p
    L.crt_oracle_scene_create.argtypes = [_p, C.c_int64, _p, C.c_int64, _p, _p, C.c_int64, _p, _p]
    L.crt_oracle_scene_create_f64.restype = _p
    L.crt_oracle_scene_create_f64.argtypes = [_p, C.c_int64, _p, C.c_int64, _p, _p, C.c_int64, _p, _p]
    L.crt_oracle_scene_destroy.argtypes = [_p]
    L.crt_oracle_tri_intersect.restype = C.c_int
    L.crt_oracle_tri_intersect.argtypes = [_p, _p, _p, _p, _p, C.c_double, C.c_double, _p]
    L.crt_oracle_closest_brute.restype = C.c_int64
    L.crt_oracle_closest_brute.argtypes = [_p, _p, _p, C.c_double, C.c_double, _p]
    L.crt_oracle_trace.argtypes = [_p, C.c_int64, _p, _p, _p, _p, C.c_int, _p, _p, _p, _p]
    L.crt_oracle_tie_flags.argtypes = [_p, C.c_int64, _p, _p, C.c_double, C.c_double, _p]
    L.crt_oracle_edge_flags.argtypes = [_p, C.c_int64, _p, _p, _p, C.c_double, C.c_double, _p]
    L.crt_oracle_camera_rays.argtypes = [_p, _p, _p]
    L.crt_oracle_pass.argtypes = [_p, _p, _p, _p, _p, _p, _p]
    L.crt_oracle_lidar_batch.argtypes = [_p, C.c_int64, _p, _p, C.c_int, _p]
    L.crt_oracle_u01.restype = C.c_double
    L.crt_oracle_u01.argtypes = [C.c_uint64] * 4
    L.crt_oracle_render.argtypes = [_p, _p, _p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64,
                                    C.c_uint64, _p, _p]
    L.crt_oracle_render_ex.argtypes = [_p, _p, _p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64,
                                       C.c_uint64, C.c_double, C.c_double, _p, _p, _p]
    L.crt_oracle_closest_fragile.argtypes = [_p, C.c_int64, _p, _p, C.c_double, _p]
    L.crt_oracle_shadow_fragile.argtypes = [_p, C.c_int64, _p, _p, _p, C.c_double, _p]
    L.crt_oracle_scene_set_materials.argtypes = [_p, C.c_int64, _p, _p, C.c_int64, C.c_int, _p]
    L.crt_oracle_radiation_pressure.argtypes = [_p, _p, C.c_double, C.c_double, C.c_double, C.c_int64, _p,
                                                C.c_double, _p, _p, _p, _p]
    L.crt_oracle_vertex_normals.argtypes = [_p, C.c_int64, _p, C.c_int64, _p]
    L.crt_oracle_num_threads.restype = C.c_int


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_camera(cam: dict) -> Camera:
    c = Camera()
    c.position[:] = [float(x) for x in np.asarray(cam["position"], dtype=np.float64)]
    c.rotation[:] = [float(x) for x in np.asarray(cam["rotation"], dtype=np.float64).reshape(9)]
    c.focal_length = float(cam["focal_length"])
    c.resolution[:] = [float(x) for x in cam["resolution"]]
    c.sensor_size[:] = [float(x) for x in cam["sensor_size"]]
    c.z_positive = int(bool(cam.get("z_positive", False)))
    return c


def make_light(light: dict) -> Light:
    l = Light()
    kind = light.get("type", "area")
    l.type = {"point": 0, "area": 1, "directional": 2}[kind]
    l.intensity = float(light["intensity"])
    l.position[:] = [float(x) for x in np.asarray(light.get("position", np.zeros(3)), dtype=np.float64)]
    l.rotation[:] = [float(x) for x in np.asarray(light.get("rotation", np.eye(3)), dtype=np.float64).reshape(9)]
    l.size[:] = [float(x) for x in light.get("size", [0.0, 0.0])]
    l.direction[:] = [float(x) for x in np.asarray(light.get("direction", np.zeros(3)), dtype=np.float64)]
    return l


class Scene:
    """FP64 oracle scene (median-split BVH over the same triangles)."""

    def __init__(self, vertices, indices, tri_inst=None, vnormals=None, colors=None, smooth=None, f64=False):
        self.v = np.ascontiguousarray(vertices, dtype=np.float64 if f64 else np.float32)
        self.i = np.ascontiguousarray(indices, dtype=np.uint32)
        self.inst = None if tri_inst is None else np.ascontiguousarray(tri_inst, dtype=np.uint32)
        self.vn = None if vnormals is None else np.ascontiguousarray(vnormals, dtype=np.float32)
        ni = 1 if self.inst is None else int(self.inst.max())
        self.colors = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 3)
        if self.colors is not None:
            ni = max(ni, self.colors.shape[0])
        self.smooth = None if smooth is None else np.ascontiguousarray(smooth, dtype=np.uint8)
        self.ni = ni
        create = lib().crt_oracle_scene_create_f64 if f64 else lib().crt_oracle_scene_create
        self.h = create(_ptr(self.v), self.v.shape[0], _ptr(self.i), self.i.shape[0], _ptr(self.inst),
                        _ptr(self.vn), ni, _ptr(self.colors), _ptr(self.smooth))
        if not self.h:
            raise ValueError("oracle scene creation failed (empty mesh?)")

    def __del__(self):
        if getattr(self, "h", None):
            lib().crt_oracle_scene_destroy(self.h)
            self.h = None

    def trace(self, o, d, tmin=None, tmax=None, any_hit=False):
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        n = o.shape[0]
        t = np.empty(n)
        tri = np.empty(n, dtype=np.int32)
        u = np.empty(n)
        v = np.empty(n)
        tmin = None if tmin is None else np.ascontiguousarray(tmin, dtype=np.float64)
        tmax = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float64)
        lib().crt_oracle_trace(self.h, n, _ptr(o), _ptr(d), _ptr(tmin), _ptr(tmax), int(any_hit), _ptr(t),
                               _ptr(tri), _ptr(u), _ptr(v))
        return t, tri, u, v

    def closest_brute(self, o, d, tmin=0.0, tmax=np.inf):
        out = np.zeros(3)
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        tri = lib().crt_oracle_closest_brute(self.h, _ptr(o), _ptr(d), tmin, tmax, _ptr(out))
        return int(tri), out

    def tie_flags(self, o, d, rel_t=1e-6, bary_eps=1e-5):
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        f = np.zeros(o.shape[0], dtype=np.uint8)
        lib().crt_oracle_tie_flags(self.h, o.shape[0], _ptr(o), _ptr(d), rel_t, bary_eps, _ptr(f))
        return f.astype(bool)

    def edge_flags(self, o, d, other_tri, rel_t=1e-6, bary_eps=1e-5):
        """Classify disagreements: True where the ray is an exact tie or grazes
        an edge/vertex of the oracle's hit or of `other_tri` (another tracer's pick)."""
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        other = np.ascontiguousarray(other_tri, dtype=np.int32)
        f = np.zeros(o.shape[0], dtype=np.uint8)
        lib().crt_oracle_edge_flags(self.h, o.shape[0], _ptr(o), _ptr(d), _ptr(other), rel_t, bary_eps, _ptr(f))
        return f.astype(bool)

    def closest_fragile(self, o, d, eps):
        """Per ray: the closest-hit decision lies within eps (distance, after
        the 1/|cos| grazing stretch) of a triangle edge or of another triangle."""
        o = np.ascontiguousarray(o, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        f = np.zeros(o.shape[0], dtype=np.uint8)
        lib().crt_oracle_closest_fragile(self.h, o.shape[0], _ptr(o), _ptr(d), float(eps), _ptr(f))
        return f.astype(bool)

    def shadow_fragile(self, o, d, eps, tmax=None):
        """Per shadow ray over [0, tmax]: occlusion differs between triangles
        grown and shrunk by eps (a terminator or an edge within eps)."""
        o = np.ascontiguousarray(o, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        tm = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float64)
        f = np.zeros(o.shape[0], dtype=np.uint8)
        lib().crt_oracle_shadow_fragile(self.h, o.shape[0], _ptr(o), _ptr(d), _ptr(tm), float(eps), _ptr(f))
        return f.astype(bool)

    def camera_pass(self, cam: dict):
        c = make_camera(cam)
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        xyz = np.zeros((h, w, 3))
        nrm = np.zeros((h, w, 3))
        inst = np.zeros((h, w), dtype=np.uint32)
        depth = np.zeros((h, w))
        tri = np.zeros((h, w), dtype=np.int32)
        lib().crt_oracle_pass(self.h, C.byref(c), _ptr(xyz), _ptr(nrm), _ptr(inst), _ptr(depth), _ptr(tri))
        return {"xyz": xyz, "normal": nrm, "instance": inst, "depth": depth, "tri": tri}

    def lidar_batch(self, pos, rot, z_positive=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
        out = np.zeros(pos.shape[0])
        lib().crt_oracle_lidar_batch(self.h, pos.shape[0], _ptr(pos), _ptr(rot), int(z_positive), _ptr(out))
        return out

    def set_materials(self, materials, uv=None, textures=()):
        """Per-instance materials (dicts for make_material), per-vertex UVs
        (V, 2) and RGBA8 textures (H, W, 4) referenced by index."""
        mats = (Material * max(len(materials), 1))(*[make_material(m) for m in materials])
        self._tex_keep = [np.ascontiguousarray(t, dtype=np.uint8) for t in textures]
        tex = (Texture * max(len(self._tex_keep), 1))(
            *[Texture(t.shape[1], t.shape[0], t.ctypes.data) for t in self._tex_keep])
        uvs = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 2)
        lib().crt_oracle_scene_set_materials(self.h, len(materials), mats, _ptr(uvs), 0 if uvs is None else
                                             uvs.shape[0], len(self._tex_keep), tex)

    def render(self, cam: dict, lights, min_samples, max_samples, noise_threshold, num_bounces, seed=0,
               key_offset=0, offset=1e-4, eps=None):
        """Radiance (H, W, 3) f32 and samples used (H, W).  `offset` is the
        self-intersection offset (the reference's 1e-4 by default); with `eps`
        also the per-pixel fragile-decision counts (H, W, 3): closest hits,
        shadow rays, texel lookups decided within eps (see crt_oracle.h)."""
        c = make_camera(cam)
        arr = (Light * max(len(lights), 1))(*[make_light(l) for l in lights])
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        rgb = np.zeros((h, w, 3), dtype=np.float32)
        spp = np.zeros((h, w), dtype=np.int32)
        frag = None if eps is None else np.zeros((h, w, 3), dtype=np.int32)
        lib().crt_oracle_render_ex(self.h, C.byref(c), arr, len(lights), min_samples, max_samples, noise_threshold,
                                   num_bounces, seed, key_offset, float(offset), float(eps or 0.0), _ptr(rgb),
                                   _ptr(spp), _ptr(frag))
        if eps is None:
            return rgb, spp
        return rgb, spp, frag

    def radiation_pressure(self, sun, pressure, rho_s, rho_d, grid, center, radius, ref=(0.0, 0.0, 0.0)):
        sun = np.ascontiguousarray(sun, dtype=np.float64)
        center = np.ascontiguousarray(center, dtype=np.float64)
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        f = np.zeros(3)
        t = np.zeros(3)
        nh = C.c_int64(0)
        lib().crt_oracle_radiation_pressure(self.h, _ptr(sun), pressure, rho_s, rho_d, grid, _ptr(center), radius,
                                            _ptr(ref), _ptr(f), _ptr(t), C.byref(nh))
        return f, t, nh.value


def camera_rays(cam: dict):
    c = make_camera(cam)
    w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
    o = np.zeros((h * w, 3))
    d = np.zeros((h * w, 3))
    lib().crt_oracle_camera_rays(C.byref(c), _ptr(o), _ptr(d))
    return o, d


def tri_intersect(p0, p1, p2, o, d, tmin=0.0, tmax=np.inf):
    out = np.zeros(3)
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (p0, p1, p2, o, d)]
    hit = lib().crt_oracle_tri_intersect(*[_ptr(x) for x in a], tmin, tmax, _ptr(out))
    return bool(hit), out


def u01(seed, a, b, c):
    return lib().crt_oracle_u01(seed, a, b, c)


def vertex_normals(vertices, indices):
    v = np.ascontiguousarray(vertices, dtype=np.float32)
    i = np.ascontiguousarray(indices, dtype=np.uint32)
    out = np.zeros_like(v)
    lib().crt_oracle_vertex_normals(_ptr(v), v.shape[0], _ptr(i), i.shape[0], _ptr(out))
    return out


def num_threads():
    return lib().crt_oracle_num_threads()
