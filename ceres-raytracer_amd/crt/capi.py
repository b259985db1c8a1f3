"""ctypes binding of the C ABI in include/crt_hip.h (libcrt_hip.so).

This is the thin Python view of the HIP core used by the tests and the
benchmark.  There is no fall-back: if libcrt_hip.so is missing or no GPU is
present, calls raise instead of computing anything on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libcrt_hip.so")

CRT_OK = 0
CRT_ERR_INVALID = 1
CRT_ERR_TYPE = 2
CRT_ERR_OOM = 3
CRT_ERR_HIP = 4
CRT_ERR_STATE = 5
CRT_ERR_ADMISSION = 6
CRT_ERR_CANCELLED = 7
CRT_ERR_TIMEOUT = 8

CRT_OUT_DEVICE = 1
CRT_ANY_HIT = 2

LIGHT_POINT, LIGHT_AREA, LIGHT_DIRECTIONAL = 0, 1, 2


class CrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("crt error %d: %s" % (code, msg))
        self.code = code


class Camera(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("rotation", C.c_double * 9), ("focal_length", C.c_double),
                ("resolution", C.c_double * 2), ("sensor_size", C.c_double * 2), ("z_positive", C.c_int32),
                ("pad_", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad_", C.c_int32), ("intensity", C.c_double), ("position", C.c_double * 3),
                ("rotation", C.c_double * 9), ("size", C.c_double * 2), ("direction", C.c_double * 3)]


class BuildOpts(C.Structure):
    _fields_ = [("max_leaf_size", C.c_uint32), ("flags", C.c_uint32)]


class BvhInfo(C.Structure):
    _fields_ = [("n_tris", C.c_uint32), ("n_wide_nodes", C.c_uint32), ("n_leaves", C.c_uint32),
                ("max_depth", C.c_uint32), ("stack_capacity", C.c_uint32), ("pad_", C.c_uint32),
                ("build_ms", C.c_double), ("sah_cost", C.c_double), ("bounds", C.c_double * 6),
                ("sphere", C.c_double * 4), ("device_bytes", C.c_uint64), ("self_offset", C.c_double),
                ("refit_ms", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("trace_ms", C.c_double), ("shadow_ms", C.c_double),
                ("rays_traced", C.c_uint64), ("shadow_rays", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RenderOpts(C.Structure):
    _fields_ = [("min_samples", C.c_int32), ("max_samples", C.c_int32), ("noise_threshold", C.c_double),
                ("num_bounces", C.c_int32), ("pad_", C.c_uint32), ("seed", C.c_uint64)]


class FlashOpts(C.Structure):
    _fields_ = [("beams", Camera), ("sub_rays", C.c_uint32), ("pad_", C.c_uint32), ("sigma", C.c_double),
                ("seed", C.c_uint64)]


class BatchOpts(C.Structure):
    _fields_ = [("samples_per_pass", C.c_uint32), ("use_graph", C.c_uint32), ("persistent", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("grab", C.c_uint32), ("extend_mode", C.c_uint32),
                ("separate_kernels", C.c_uint32), ("coherence_sort", C.c_uint32), ("rays_per_octet", C.c_uint32),
                ("rays_per_quad", C.c_uint32), ("tile_width", C.c_uint32), ("fixed_share", C.c_uint32)]


class FluxOpts(C.Structure):
    _fields_ = [("sun_dir", C.c_double * 3), ("pressure", C.c_double), ("rho_s", C.c_double),
                ("rho_d", C.c_double), ("grid", C.c_uint32), ("pad_", C.c_uint32), ("center", C.c_double * 3),
                ("radius", C.c_double), ("ref_point", C.c_double * 3)]


class ServerOpts(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("n_workers", C.c_uint32), ("max_batch", C.c_uint32),
                ("max_batch_paths", C.c_uint64), ("batching", C.c_uint32), ("pad_", C.c_uint32),
                ("render", BatchOpts)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("scene", C.c_uint32), ("camera", Camera), ("spp", C.c_int32),
                ("bounces", C.c_int32), ("seed", C.c_uint64), ("n_lights", C.c_uint32), ("sub_rays", C.c_uint32),
                ("lights", Light * 2), ("sigma", C.c_double), ("flux", FluxOpts),
                ("rgba", C.c_void_p), ("lidar", C.c_void_p), ("force_torque", C.c_void_p),
                ("light_list", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("rejected", C.c_uint64), ("rays", C.c_uint64), ("batches", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("wall_s", C.c_double),
                ("mrays_per_s", C.c_double), ("cancelled", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Material(C.Structure):
    """crt_material: src/crt/materials/material.hpp:90-184."""
    _fields_ = [("type", C.c_uint32), ("texture", C.c_int32), ("spec_texture", C.c_int32), ("alpha", C.c_float),
                ("color", C.c_float * 3), ("coeffs", C.c_float * 3)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p)]


MATERIAL_TYPES = {"lambertian": 0, "textured_lambertian": 1, "blinn_phong": 2, "mirror": 3}


def make_material(m: dict) -> Material:
    """dict(type=lambertian|textured_lambertian|blinn_phong|mirror, color=(r,g,b),
    texture=index or -1, spec_texture=index or -1, coeffs=(ka,kd,ks), alpha=24)."""
    x = Material()
    x.type = MATERIAL_TYPES[m.get("type", "lambertian")]
    x.texture = int(m.get("texture", -1))
    x.spec_texture = int(m.get("spec_texture", -1))
    x.alpha = float(m.get("alpha", 24.0))
    x.color[:] = [float(c) for c in m.get("color", (1.0, 1.0, 1.0))]
    x.coeffs[:] = [float(c) for c in m.get("coeffs", (0.0, 1.0, 0.0))]
    return x


REQ_RENDER, REQ_LIDAR, REQ_FLUX = 0, 1, 2

_p = C.c_void_p
_lib = None

# name -> (restype, argtypes); the exported-symbol test checks this list
# against include/crt_hip.h.
SIGNATURES = {
    "crt_device_init": (C.c_int, [C.c_int]),
    "crt_last_error": (C.c_char_p, []),
    "crt_version": (C.c_int, []),
    "crt_timing_log_enable": (C.c_int, [C.c_int]),
    "crt_timing_log_read": (C.c_int, [C.c_char_p, _p, _p]),
    "crt_scene_create": (C.c_int, [_p, C.c_uint32, _p, C.c_uint32, _p, _p, C.c_uint32, _p, _p, _p, _p]),
    "crt_scene_create_f64": (C.c_int, [_p, C.c_uint32, _p, C.c_uint32, _p, _p, C.c_uint32, _p, _p, _p, _p]),
    "crt_scene_info": (C.c_int, [_p, _p]),
    "crt_scene_set_materials": (C.c_int, [_p, C.c_uint32, _p, _p, C.c_uint32, _p]),
    "crt_scene_refit": (C.c_int, [_p, _p, _p, C.c_double, _p]),
    "crt_scene_destroy": (None, [_p]),
    "crt_scene_serialize": (C.c_int, [_p, _p, _p]),
    "crt_scene_deserialize": (C.c_int, [_p, C.c_uint64, _p]),
    "crt_pass": (C.c_int, [_p, _p, _p, _p, _p, _p, _p, C.c_uint32, _p, _p]),
    "crt_trace_rays": (C.c_int, [_p, C.c_uint64, _p, _p, _p, _p, _p, _p, _p, _p, C.c_uint32, _p, _p]),
    "crt_trace_stats": (C.c_int, [_p, C.c_uint64, _p, _p, C.c_uint32, _p, _p, _p, _p]),
    "crt_sun_pass": (C.c_int, [_p, _p, _p, _p, _p, _p, C.c_uint32, _p, _p]),
    "crt_render": (C.c_int, [_p, _p, _p, C.c_uint32, _p, _p, _p, _p, C.c_uint32, _p, _p]),
    "crt_render_batch": (C.c_int, [_p, C.c_uint32, _p, _p, _p, _p, C.c_uint32, _p, _p, _p, _p, _p, C.c_uint32,
                                   _p, _p]),
    "crt_server_create": (C.c_int, [_p, _p]),
    "crt_server_add_scene": (C.c_int, [_p, _p, _p]),
    "crt_server_submit": (C.c_int, [_p, _p, _p]),
    "crt_server_wait": (C.c_int, [_p, C.c_uint64, _p]),
    "crt_server_wait_for": (C.c_int, [_p, C.c_uint64, C.c_double, _p]),
    "crt_server_cancel": (C.c_int, [_p, C.c_uint64]),
    "crt_server_drain": (C.c_int, [_p]),
    "crt_server_stats_get": (C.c_int, [_p, _p]),
    "crt_server_stats_reset": (C.c_int, [_p]),
    "crt_server_destroy": (None, [_p]),
    "crt_lidar_batch": (C.c_int, [_p, C.c_uint64, _p, _p, C.c_int32, _p, C.c_uint32, _p, _p]),
    "crt_flash_lidar": (C.c_int, [_p, _p, _p, _p, _p, _p, C.c_uint32, _p, _p]),
    "crt_radiation_pressure": (C.c_int, [_p, _p, _p, _p, _p, _p, _p]),
    "crt_radiation_pressure_noalloc": (C.c_int, [_p, _p, _p, _p]),
}


def load(path: str = LIB_PATH):
    """Load libcrt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("libcrt_hip.so not found at %s: run __graft_entry__.build()" % path)
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


BUILD_PLOC = 1   # crt_build_opts.flags: PLOC clustering whatever the size
BUILD_LBVH = 4   # crt_build_opts.flags: Morton-order LBVH whatever the size (default: PLOC up to 131072 triangles)
BUILD_GREEDY_WIDE = 2   # crt_build_opts.flags: greedy collapse into 8-wide nodes instead of the SAH-optimal DP


def _check(code):
    if code != CRT_OK:
        msg = load().crt_last_error().decode(errors="replace")
        if code == CRT_ERR_INVALID:
            raise ValueError(msg)
        if code == CRT_ERR_TYPE:
            raise TypeError(msg)
        if code == CRT_ERR_OOM:
            raise MemoryError(msg)
        if code == CRT_ERR_TIMEOUT:
            raise TimeoutError(msg)
        raise CrtError(code, msg)


_initialised = False


def init(device: int = 0):
    global _initialised
    if not _initialised:
        _check(load().crt_device_init(device))
        _initialised = True


def ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(int(a))   # raw device address (e.g. DeviceBuffer.ptr)


def _hip():
    """The HIP runtime libcrt_hip.so is bound to (symbols looked up through
    the library's handle), so the device buffers and synchronisation below
    are those of the runtime its entry points use."""
    L = load()
    if not getattr(L, "_hip_ready", False):
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L._hip_ready = True
    return L


class Stream:
    """A HIP stream (default priority) of the library's runtime; `.handle`
    is the hipStream_t to pass as `stream`.  Calls on one scene with
    different streams use separate workspaces, so their device work may
    overlap."""

    def __init__(self):
        p = C.c_void_p()
        if _hip().hipStreamCreate(C.byref(p)) != 0:
            raise RuntimeError("hipStreamCreate failed")
        self._s = p

    @property
    def handle(self) -> int:
        return self._s.value

    def synchronize(self):
        if _hip().hipStreamSynchronize(self._s) != 0:
            raise RuntimeError("hipStreamSynchronize failed")

    def destroy(self):
        if self._s.value:
            _hip().hipStreamDestroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory for CRT_OUT_DEVICE outputs (hipMalloc of the library's
    runtime).  `.ptr` is the device address; `.to_host(array)` copies back."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        if _hip().hipMalloc(C.byref(p), C.c_size_t(max(self.nbytes, 1))) != 0:
            raise MemoryError("hipMalloc of %d bytes failed" % self.nbytes)
        self._p = p

    @property
    def ptr(self) -> int:
        return self._p.value

    def fill(self, byte: int):
        if _hip().hipMemset(self._p, int(byte), C.c_size_t(self.nbytes)) != 0:
            raise RuntimeError("hipMemset failed")

    def from_host(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if _hip().hipMemcpy(self._p, a.ctypes.data, C.c_size_t(a.nbytes), 1) != 0:   # host to device
            raise RuntimeError("hipMemcpy failed")
        return self

    def to_host(self, out: np.ndarray) -> np.ndarray:
        assert out.nbytes <= self.nbytes and out.flags["C_CONTIGUOUS"]
        if _hip().hipMemcpy(out.ctypes.data, self._p, C.c_size_t(out.nbytes), 2) != 0:   # device to host
            raise RuntimeError("hipMemcpy failed")
        return out

    def free(self):
        if self._p.value:
            _hip().hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_copy(dst: DeviceBuffer, src: DeviceBuffer):
    """Device-to-device copy of src into dst (the runtime's blit kernel)."""
    n = min(dst.nbytes, src.nbytes)
    if _hip().hipMemcpy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), C.c_size_t(n), 3) != 0:
        raise RuntimeError("hipMemcpy failed")


def timing_log(on: bool):
    """Enable / disable the library's kernel timing log (crt_timing_log_enable)."""
    _check(load().crt_timing_log_enable(int(bool(on))))


def timing_log_read(kernel: str):
    """(count, mean ms) of the logged launches of `kernel` (events waited for)."""
    n = C.c_uint64()
    ms = C.c_double()
    _check(load().crt_timing_log_read(kernel.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value


def device_synchronize():
    """Wait for all work the library has enqueued on the device."""
    if _hip().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def make_camera(cam: dict) -> Camera:
    c = Camera()
    c.position[:] = [float(x) for x in np.asarray(cam["position"], dtype=np.float64).reshape(3)]
    c.rotation[:] = [float(x) for x in np.asarray(cam["rotation"], dtype=np.float64).reshape(9)]
    c.focal_length = float(cam["focal_length"])
    c.resolution[:] = [float(x) for x in cam["resolution"]]
    c.sensor_size[:] = [float(x) for x in cam["sensor_size"]]
    c.z_positive = int(bool(cam.get("z_positive", False)))
    return c


def make_light(light: dict) -> Light:
    l = Light()
    l.type = {"point": LIGHT_POINT, "area": LIGHT_AREA, "directional": LIGHT_DIRECTIONAL}[light.get("type", "area")]
    l.intensity = float(light["intensity"])
    l.position[:] = [float(x) for x in np.asarray(light.get("position", np.zeros(3)), dtype=np.float64)]
    l.rotation[:] = [float(x) for x in np.asarray(light.get("rotation", np.eye(3)), dtype=np.float64).reshape(9)]
    l.size[:] = [float(x) for x in light.get("size", [0.0, 0.0])]
    l.direction[:] = [float(x) for x in np.asarray(light.get("direction", np.zeros(3)), dtype=np.float64)]
    return l


class Scene:
    """A mesh + GPU-built BVH resident in HBM (crt_scene_create)."""

    def __init__(self, vertices, indices, tri_instance=None, vertex_normals=None, colors=None, smooth=None,
                 max_leaf_size=8, ploc=None, greedy_wide=False, f64=False):
        """f64=True hands the vertices over in FP64 (crt_scene_create_f64):
        the frame origin is subtracted before the rounding to FP32.  ploc:
        None = the library's choice by size, True = PLOC, False = LBVH."""
        init()
        v = np.ascontiguousarray(vertices, dtype=np.float64 if f64 else np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        inst = None if tri_instance is None else np.ascontiguousarray(tri_instance, dtype=np.uint32)
        vn = None if vertex_normals is None else np.ascontiguousarray(vertex_normals, dtype=np.float32)
        ni = 1 if inst is None else int(inst.max())
        col = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 3)
        if col is not None:
            ni = max(ni, col.shape[0])
        sm = None if smooth is None else np.ascontiguousarray(smooth, dtype=np.uint8)
        tree = 0 if ploc is None else (BUILD_PLOC if ploc else BUILD_LBVH)
        opts = BuildOpts(max_leaf_size, tree | (BUILD_GREEDY_WIDE if greedy_wide else 0))
        h = C.c_void_p()
        create = load().crt_scene_create_f64 if f64 else load().crt_scene_create
        _check(create(ptr(v), v.shape[0], ptr(i), i.shape[0], ptr(inst), ptr(vn), ni, ptr(col), ptr(sm),
                      C.byref(opts), C.byref(h)))
        self.h = h
        self.n_tris = i.shape[0]

    def set_materials(self, materials, uv=None, textures=()):
        """One material dict per instance (see make_material), per-vertex UVs
        (V, 2) and RGBA8 textures (H, W, 4) referenced by index
        (crt_scene_set_materials)."""
        mats = (Material * max(len(materials), 1))(*[make_material(m) for m in materials])
        keep = [np.ascontiguousarray(t, dtype=np.uint8) for t in textures]
        for t in keep:
            if t.ndim != 3 or t.shape[2] != 4:
                raise ValueError("textures must be (H, W, 4) RGBA8 arrays")
        tex = (Texture * max(len(keep), 1))(*[Texture(t.shape[1], t.shape[0], t.ctypes.data) for t in keep])
        uvs = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 2)
        _check(load().crt_scene_set_materials(self.h, len(materials), mats, ptr(uvs), len(keep), tex))

    def refit(self, vertices, vertex_normals=None, max_sah_growth=0.0) -> bool:
        """Move the vertices (same triangles) and refit the BVH on the GPU;
        returns True when the SAH had grown past max_sah_growth (default
        1.5) and the BVH was rebuilt instead (crt_scene_refit)."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        vn = None if vertex_normals is None else np.ascontiguousarray(vertex_normals, dtype=np.float32)
        rb = C.c_uint32(0)
        _check(load().crt_scene_refit(self.h, ptr(v), ptr(vn), float(max_sah_growth), C.byref(rb)))
        return bool(rb.value)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().crt_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- blobs: a built scene saved once, loaded without a rebuild --------
    def to_bytes(self) -> bytes:
        size = C.c_uint64(0)
        _check(load().crt_scene_serialize(self.h, None, C.byref(size)))
        buf = np.zeros(size.value, dtype=np.uint8)
        _check(load().crt_scene_serialize(self.h, ptr(buf), C.byref(size)))
        return buf[:size.value].tobytes()

    def save(self, path):
        with open(path, "wb") as fh:
            fh.write(self.to_bytes())

    @classmethod
    def from_bytes(cls, blob) -> "Scene":
        init()
        buf = np.frombuffer(blob, dtype=np.uint8)
        buf = np.ascontiguousarray(buf)
        h = C.c_void_p()
        _check(load().crt_scene_deserialize(ptr(buf), buf.size, C.byref(h)))
        obj = cls.__new__(cls)
        obj.h = h
        obj.n_tris = obj.info()["n_tris"]
        return obj

    @classmethod
    def load(cls, path) -> "Scene":
        with open(path, "rb") as fh:
            return cls.from_bytes(fh.read())

    def info(self) -> dict:
        inf = BvhInfo()
        _check(load().crt_scene_info(self.h, C.byref(inf)))
        d = {k: getattr(inf, k) for k, _ in BvhInfo._fields_ if k != "pad_"}
        d["bounds"] = list(inf.bounds)
        d["sphere"] = list(inf.sphere)
        return d

    def camera_pass(self, cam: dict, want=("xyz", "normal", "instance", "depth", "tri")):
        c = make_camera(cam)
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        out = {}
        if "xyz" in want: out["xyz"] = np.zeros((h, w, 3))
        if "normal" in want: out["normal"] = np.zeros((h, w, 3))
        if "instance" in want: out["instance"] = np.zeros((h, w), dtype=np.uint32)
        if "depth" in want: out["depth"] = np.zeros((h, w))
        if "tri" in want: out["tri"] = np.zeros((h, w), dtype=np.int32)
        tm = Timing()
        _check(load().crt_pass(self.h, C.byref(c), ptr(out.get("xyz")), ptr(out.get("normal")),
                               ptr(out.get("instance")), ptr(out.get("depth")), ptr(out.get("tri")), 0, None,
                               C.byref(tm)))
        out["timing"] = tm.as_dict()
        return out

    def trace(self, o, d, tmin=None, tmax=None, any_hit=False):
        o = np.ascontiguousarray(o, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        t = np.empty(n)
        tri = np.empty(n, dtype=np.int32)
        u = np.empty(n)
        v = np.empty(n)
        tmin = None if tmin is None else np.ascontiguousarray(tmin, dtype=np.float64)
        tmax = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float64)
        tm = Timing()
        flags = CRT_ANY_HIT if any_hit else 0
        _check(load().crt_trace_rays(self.h, n, ptr(o), ptr(d), ptr(tmin), ptr(tmax), ptr(t), ptr(tri), ptr(u),
                                     ptr(v), flags, None, C.byref(tm)))
        return t, tri, u, v

    def trace_stats(self, o, d, any_hit=False):
        """Traversal statistics per ray (crt_trace_stats): dict of `nodes`,
        `boxes`, `tris` (uint32, n) and the hit `t`, `tri` of the same rays."""
        o = np.ascontiguousarray(o, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        counts = np.empty((n, 3), dtype=np.uint32)
        t = np.empty(n)
        tri = np.empty(n, dtype=np.int32)
        _check(load().crt_trace_stats(self.h, n, ptr(o), ptr(d), CRT_ANY_HIT if any_hit else 0, ptr(counts), ptr(t),
                                      ptr(tri), None))
        return {"nodes": counts[:, 0], "boxes": counts[:, 1], "tris": counts[:, 2], "t": t, "tri": tri}

    def sun_pass(self, cam: dict, sun_dir, outputs=None, device=False, stream=None, timing=True):
        """Primary closest hit + one sun shadow ray per hit.  With device=True,
        `outputs` holds device addresses {depth, tri, lit} (ints) and nothing is
        copied back; with timing=False as well the call only enqueues the work
        on `stream` (no synchronisation) and returns None."""
        c = make_camera(cam)
        s = (C.c_double * 3)(*[float(x) for x in sun_dir])
        tm = Timing()
        if device:
            o = outputs or {}
            _check(load().crt_sun_pass(self.h, C.byref(c), s, ptr(o.get("depth")), ptr(o.get("tri")),
                                       ptr(o.get("lit")), CRT_OUT_DEVICE, stream, C.byref(tm) if timing else None))
            return tm.as_dict() if timing else None
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        depth = np.zeros((h, w))
        tri = np.zeros((h, w), dtype=np.int32)
        lit = np.zeros((h, w), dtype=np.uint8)
        _check(load().crt_sun_pass(self.h, C.byref(c), s, ptr(depth), ptr(tri), ptr(lit), 0, stream, C.byref(tm)))
        return {"depth": depth, "tri": tri, "lit": lit, "timing": tm.as_dict()}

    # ---- render / sensors -------------------------------------------------
    def render(self, cam: dict, lights, min_samples=1, max_samples=1, noise_threshold=1.0, num_bounces=1, seed=0):
        """Path-traced radiance: returns (rgba f32 (H,W,4), rgba uint8 (H,W,4), spp_used, timing)."""
        c = make_camera(cam)
        arr = (Light * max(len(lights), 1))(*[make_light(l) for l in lights])
        opts = RenderOpts(int(min_samples), int(max_samples), float(noise_threshold), int(num_bounces), 0, int(seed))
        w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
        rgba = np.zeros((h, w, 4), dtype=np.float32)
        u8 = np.zeros((h, w, 4), dtype=np.uint8)
        spp = np.zeros((h, w), dtype=np.int32)
        tm = Timing()
        _check(load().crt_render(self.h, C.byref(c), arr, len(lights), C.byref(opts), ptr(rgba), ptr(u8), ptr(spp),
                                 0, None, C.byref(tm)))
        return rgba, u8, spp, tm.as_dict()

    def lidar_batch(self, positions, rotations, z_positive=False):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        rot = np.ascontiguousarray(rotations, dtype=np.float64).reshape(-1, 9)
        out = np.zeros(pos.shape[0])
        tm = Timing()
        _check(load().crt_lidar_batch(self.h, pos.shape[0], ptr(pos), ptr(rot), int(bool(z_positive)), ptr(out), 0,
                                      None, C.byref(tm)))
        return out

    def lidar_batch_device(self, n, positions, rotations, ranges, z_positive=False, stream=None):
        """crt_lidar_batch on device-resident poses: `positions` (n x 3 f64),
        `rotations` (n x 9 f64, row-major R per pose) and `ranges` (n f64) are
        device addresses; only enqueues the work on `stream`."""
        _check(load().crt_lidar_batch(self.h, int(n), ptr(positions), ptr(rotations), int(bool(z_positive)),
                                      ptr(ranges), CRT_OUT_DEVICE, stream, None))

    def flash_lidar(self, beams: dict, sub_rays=8, sigma=5e-4, seed=0, outputs=None, device=False, stream=None,
                    timing=True):
        """Flash LiDAR over a beam grid.  With device=True, `outputs` holds
        device addresses {single, first, last, n_returns} (ints) and only the
        timing is returned; with timing=False as well the call only enqueues
        the work on `stream` and returns None."""
        o = FlashOpts()
        o.beams = make_camera(beams)
        o.sub_rays = int(sub_rays)
        o.sigma = float(sigma)
        o.seed = int(seed)
        if device:
            d = outputs or {}
            tm = Timing()
            _check(load().crt_flash_lidar(self.h, C.byref(o), ptr(d.get("single")), ptr(d.get("first")),
                                          ptr(d.get("last")), ptr(d.get("n_returns")), CRT_OUT_DEVICE, stream,
                                          C.byref(tm) if timing else None))
            return tm.as_dict() if timing else None
        w, h = int(np.floor(beams["resolution"][0])), int(np.floor(beams["resolution"][1]))
        single, first, last = np.zeros((h, w)), np.zeros((h, w)), np.zeros((h, w))
        nret = np.zeros((h, w), dtype=np.uint8)
        tm = Timing()
        _check(load().crt_flash_lidar(self.h, C.byref(o), ptr(single), ptr(first), ptr(last), ptr(nret), 0, None,
                                      C.byref(tm)))
        return {"single": single, "first": first, "last": last, "n_returns": nret, "timing": tm.as_dict()}

    def radiation_pressure(self, sun_dir, pressure, rho_s, rho_d, grid, center=(0, 0, 0), radius=0.0,
                           ref_point=(0, 0, 0), stream=None):
        o = FluxOpts()
        o.sun_dir[:] = [float(x) for x in sun_dir]
        o.pressure, o.rho_s, o.rho_d, o.grid = float(pressure), float(rho_s), float(rho_d), int(grid)
        o.center[:] = [float(x) for x in center]
        o.radius = float(radius)
        o.ref_point[:] = [float(x) for x in ref_point]
        f = (C.c_double * 3)()
        t = (C.c_double * 3)()
        nh = C.c_uint64(0)
        tm = Timing()
        _check(load().crt_radiation_pressure(self.h, C.byref(o), f, t, C.byref(nh), stream, C.byref(tm)))
        return np.array(f[:]), np.array(t[:]), nh.value, tm.as_dict()

    def render_batch(self, cams, lights, min_samples=1, max_samples=1, noise_threshold=0.0, num_bounces=1,
                     seeds=None, spp=None, samples_per_pass=0, use_graph=False, persistent=False, blocks_per_cu=0,
                     grab=0, extend_mode=0, separate_kernels=0, coherence_sort=0, rays_per_octet=0, rays_per_quad=0,
                     tile_width=0, fixed_share=0, outputs=True, stream=None):
        """Render many cameras in one set of launches.  Returns a list of
        per-camera (rgba f32, spp_used) and the timing dict."""
        n = len(cams)
        arr = (Camera * n)(*[make_camera(c) for c in cams])
        larr = (Light * max(len(lights), 1))(*[make_light(l) for l in lights])
        opts = RenderOpts(int(min_samples), int(max_samples), float(noise_threshold), int(num_bounces), 0, 0)
        bo = BatchOpts(int(samples_per_pass), int(bool(use_graph)), int(bool(persistent)), int(blocks_per_cu),
                       int(grab), int(extend_mode), int(separate_kernels), int(coherence_sort), int(rays_per_octet),
                       int(rays_per_quad), int(tile_width), int(fixed_share))
        sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint64)
        ms = None if spp is None else np.ascontiguousarray(spp, dtype=np.int32)
        sizes = [(int(np.floor(c["resolution"][1])), int(np.floor(c["resolution"][0]))) for c in cams]
        total = sum(h * w for h, w in sizes)
        rgba = np.zeros((total, 4), dtype=np.float32) if outputs else None
        sppo = np.zeros(total, dtype=np.int32) if outputs else None
        tm = Timing()
        _check(load().crt_render_batch(self.h, n, arr, ptr(ms), ptr(sd), larr, len(lights), C.byref(opts),
                                       C.byref(bo), ptr(rgba), None, ptr(sppo), 0, stream, C.byref(tm)))
        out = []
        if outputs:
            off = 0
            for h, w in sizes:
                out.append((rgba[off:off + h * w].reshape(h, w, 4), sppo[off:off + h * w].reshape(h, w)))
                off += h * w
        return out, tm.as_dict()
