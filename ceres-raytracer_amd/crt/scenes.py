"""Deterministic synthetic scenes used by the tests and the benchmark.

Two families, both seeded and fully reproducible:

* ``cornell_box(seed)`` -- a 36-triangle Cornell box: 5 walls of the unit cube
  (front face open), two yawed boxes with seeded sizes, and a 0.25 x 0.25 light
  quad just below the ceiling.  Each part is its own entity (instance ids 1..6).
* ``displaced_icosphere(level, radius, ...)`` -- an icosphere subdivided
  ``level`` times (20 * 4**level faces) whose vertices are pushed radially by
  a seeded 6-octave value-noise fBm.

Meshes are returned as FP32 vertex arrays plus uint32 triangle indices, wound
counter-clockwise when seen from the visible side.  The reference tracer
computes its geometric normal as ``(p0 - p1) x (p2 - p0)``
(lib/bvh/triangle.hpp:39), i.e. pointing *away* from the visible side, and
its shading and offsets assume exactly that (src/crt/path_tracing/
unidirectional.hpp:60-62, src/crt/materials/brdfs.hpp:18-42), so the winding
matters.

``write_obj`` writes vertices with ``%.9g`` so that the reference's ``strtof``
parser (lib/model_loaders/obj.hpp:75-77) reads back exactly the FP32 values.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np


@dataclass
class Mesh:
    vertices: np.ndarray          # (V, 3) float32
    indices: np.ndarray           # (F, 3) uint32
    name: str = "mesh"

    @property
    def n_tris(self) -> int:
        return int(self.indices.shape[0])


@dataclass
class Scene:
    meshes: List[Mesh]
    colors: List[Tuple[float, float, float]]
    smooth: List[bool]
    meta: dict = field(default_factory=dict)

    def merged(self):
        """Concatenate all meshes: vertices (V,3) f32, indices (F,3) u32,
        per-triangle instance id (F,) u32 starting at 1 (list order), which is
        the id assignment of src/pybind.cpp:110-120."""
        verts, idx, inst = [], [], []
        base = 0
        for k, m in enumerate(self.meshes):
            verts.append(m.vertices)
            idx.append(m.indices.astype(np.int64) + base)
            inst.append(np.full(m.n_tris, k + 1, dtype=np.uint32))
            base += m.vertices.shape[0]
        return (np.ascontiguousarray(np.concatenate(verts).astype(np.float32)),
                np.ascontiguousarray(np.concatenate(idx).astype(np.uint32)),
                np.ascontiguousarray(np.concatenate(inst)))


def write_obj(path: str, mesh: Mesh) -> None:
    """Write ``v``/``f`` records; FP32 values round-trip exactly through strtof."""
    v = mesh.vertices.astype(np.float32)
    f = mesh.indices.astype(np.int64) + 1
    with open(path, "w") as fh:
        fh.write("# crt synthetic mesh %s\n" % mesh.name)
        fh.write("".join("v %.9g %.9g %.9g\n" % (a, b, c) for a, b, c in v.tolist()))
        fh.write("".join("f %d %d %d\n" % (a, b, c) for a, b, c in f.tolist()))


# --------------------------------------------------------------------------
# Cornell box
# --------------------------------------------------------------------------

def _quad(a, b, c, d):
    """Two triangles (a,b,c), (a,c,d) for a quad given CCW from the visible side."""
    v = np.array([a, b, c, d], dtype=np.float64)
    return v, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int64)


def _box(center_xz, size, yaw):
    """Closed box on the floor, CCW seen from outside (12 triangles)."""
    w, h, d = size
    hx, hz = w / 2.0, d / 2.0
    local = np.array([[-hx, 0, -hz], [hx, 0, -hz], [hx, 0, hz], [-hx, 0, hz],
                      [-hx, h, -hz], [hx, h, -hz], [hx, h, hz], [-hx, h, hz]])
    c, s = math.cos(yaw), math.sin(yaw)
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    pts = local @ rot.T + np.array([center_xz[0], 0.0, center_xz[1]])
    quads = [(0, 1, 2, 3),   # bottom, seen from below
             (4, 7, 6, 5),   # top, seen from above
             (0, 4, 5, 1),   # -z side
             (2, 6, 7, 3),   # +z side
             (1, 5, 6, 2),   # +x side
             (3, 7, 4, 0)]   # -x side
    faces = []
    for a, b, cc, dd in quads:
        faces += [[a, b, cc], [a, cc, dd]]
    faces = np.array(faces, dtype=np.int64)
    # orient every face CCW seen from outside: (p1-p0)x(p2-p0) along (centroid - box centre)
    centre = pts.mean(axis=0)
    for k in range(faces.shape[0]):
        p0, p1, p2 = pts[faces[k]]
        nrm = np.cross(p1 - p0, p2 - p0)
        if np.dot(nrm, (p0 + p1 + p2) / 3.0 - centre) < 0:
            faces[k] = faces[k][[0, 2, 1]]
    return pts, faces


CORNELL_GREY = (0.75, 0.75, 0.75)
CORNELL_RED = (0.75, 0.10, 0.10)
CORNELL_GREEN = (0.10, 0.75, 0.10)


def cornell_box(seed: int = 0) -> Scene:
    """36-triangle Cornell box inside the unit cube [0,1]^3, open at z=1.

    Entities (instance ids): 1 grey floor/ceiling/back (6 tris), 2 red left
    wall x=0 (2), 3 green right wall x=1 (2), 4 and 5 boxes (12 each) with
    sizes U(0.2,0.4) per axis and yaw U(0,45) deg, 6 the light quad (2) at
    y = 1 - 1e-3 facing down.
    """
    rng = np.random.default_rng(seed)
    meshes, colors = [], []
    # walls seen from inside the room -> CCW from inside
    floor = _quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0))
    ceil = _quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1))
    back = _quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    left = _quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1))
    right = _quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0))

    def merge(parts):
        vs, fs, base = [], [], 0
        for v, f in parts:
            vs.append(v)
            fs.append(f + base)
            base += v.shape[0]
        return np.concatenate(vs), np.concatenate(fs)

    def orient_inside(v, f):
        # visible side is the room interior (centre 0.5,0.5,0.5): make
        # (p1-p0)x(p2-p0) point into the room
        f = f.copy()
        for k in range(f.shape[0]):
            p0, p1, p2 = v[f[k]]
            nrm = np.cross(p1 - p0, p2 - p0)
            if np.dot(nrm, np.array([0.5, 0.5, 0.5]) - (p0 + p1 + p2) / 3.0) < 0:
                f[k] = f[k][[0, 2, 1]]
        return v, f

    for parts, col, name in [((floor, ceil, back), CORNELL_GREY, "walls"),
                             ((left,), CORNELL_RED, "red_wall"),
                             ((right,), CORNELL_GREEN, "green_wall")]:
        v, f = orient_inside(*merge(parts))
        meshes.append(Mesh(v.astype(np.float32), f.astype(np.uint32), name))
        colors.append(col)
    box_specs = []
    for k, cxz in enumerate([(0.32, 0.38), (0.68, 0.62)]):
        size = rng.uniform(0.2, 0.4, size=3)
        yaw = math.radians(rng.uniform(0.0, 45.0))
        v, f = _box(cxz, size, yaw)
        meshes.append(Mesh(v.astype(np.float32), f.astype(np.uint32), "box%d" % k))
        colors.append(CORNELL_GREY)
        box_specs.append({"center_xz": list(cxz), "size": size.tolist(), "yaw_deg": math.degrees(yaw)})
    # light quad facing down (visible from below): CCW seen from below
    y = 1.0 - 1e-3
    lq = _quad((0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625))
    v, f = lq
    f = f.copy()
    for k in range(f.shape[0]):
        p0, p1, p2 = v[f[k]]
        if np.cross(p1 - p0, p2 - p0)[1] > 0:   # must point down (towards the room)
            f[k] = f[k][[0, 2, 1]]
    meshes.append(Mesh(v.astype(np.float32), f.astype(np.uint32), "light_quad"))
    colors.append(CORNELL_GREY)
    meta = {"seed": seed, "boxes": box_specs}
    return Scene(meshes, colors, [False] * len(meshes), meta)


def cornell_camera():
    """Pinhole camera of the Cornell config: 256x256, f=1, 60 deg FOV
    (sensor = 2 f tan 30 deg), at (0.5, 0.5, 1.9) looking down -z."""
    f = 1.0
    s = 2.0 * f * math.tan(math.radians(30.0))
    return {"focal_length": f, "resolution": [256, 256], "sensor_size": [s, s],
            "z_positive": False, "position": np.array([0.5, 0.5, 1.9]),
            "rotation": np.eye(3)}


def cornell_light():
    """AreaLight of the Cornell config: 0.25 x 0.25, sample plane 2e-3 below
    the light quad, facing down; the local XY plane maps to world XZ.

    The reference places area-light samples with R (not R^T):
    src/crt/lights/area_light.hpp:63 -> src/crt/transform.hpp:31-40."""
    rot = np.array([[1.0, 0.0, 0.0],
                    [0.0, 0.0, -1.0],
                    [0.0, 1.0, 0.0]])
    return {"intensity": 0.15, "size": [0.25, 0.25],
            "position": np.array([0.5, 1.0 - 3e-3, 0.5]), "rotation": rot}


# --------------------------------------------------------------------------
# Displaced icosphere
# --------------------------------------------------------------------------

def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    # CCW seen from outside
    for k in range(f.shape[0]):
        p0, p1, p2 = v[f[k]]
        if np.dot(np.cross(p1 - p0, p2 - p0), p0 + p1 + p2) < 0:
            f[k] = f[k][[0, 2, 1]]
    return v, f


def icosphere(level: int):
    """Unit icosphere: 10*4**level + 2 vertices, 20*4**level faces, CCW outward."""
    v, f = _icosahedron()
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (1 << 32) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = (uniq >> 32).astype(np.int64)
        b = (uniq & 0xFFFFFFFF).astype(np.int64)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        m01 = nv + inv[:nf]
        m12 = nv + inv[nf:2 * nf]
        m20 = nv + inv[2 * nf:]
        f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
                            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _hash_u32(x):
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def _lattice(ix, iy, iz, seed):
    """Value in [-1,1] at integer lattice points (seeded integer hash)."""
    with np.errstate(over="ignore"):
        h = _hash_u32(ix.astype(np.int64).astype(np.uint32) * np.uint32(0x8DA6B343)
                      ^ iy.astype(np.int64).astype(np.uint32) * np.uint32(0xD8163841)
                      ^ iz.astype(np.int64).astype(np.uint32) * np.uint32(0xCB1AB31F)
                      ^ np.uint32(seed * 0x9E3779B9 & 0xFFFFFFFF))
    return h.astype(np.float64) * (2.0 / 4294967295.0) - 1.0


def value_noise(p, seed=0):
    """Trilinear value noise with smoothstep fade, range [-1,1]."""
    fl = np.floor(p)
    i = fl.astype(np.int64)
    t = p - fl
    t = t * t * (3.0 - 2.0 * t)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        wx = t[:, 0] if dx else 1.0 - t[:, 0]
        for dy in (0, 1):
            wy = t[:, 1] if dy else 1.0 - t[:, 1]
            for dz in (0, 1):
                wz = t[:, 2] if dz else 1.0 - t[:, 2]
                out += wx * wy * wz * _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def fbm(p, octaves=6, lacunarity=2.0, gain=0.5, base_frequency=2.0, seed=0):
    """Fractal sum of value noise normalised to [-1,1].  Every point is
    independent, so large inputs are split into chunks summed on threads
    (numpy releases the GIL inside its loops); the values are the same bits
    as one pass over all points."""
    chunk = 1 << 18
    if p.shape[0] > 2 * chunk:
        from concurrent.futures import ThreadPoolExecutor
        parts = [p[i:i + chunk] for i in range(0, p.shape[0], chunk)]
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1, len(parts))) as pool:
            return np.concatenate(list(pool.map(
                lambda q: fbm(q, octaves, lacunarity, gain, base_frequency, seed), parts)))
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_frequency, 0.0
    for o in range(octaves):
        total += amp * value_noise(p * freq, seed + o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def displaced_icosphere(level: int = 8, radius: float = 2000.0, amplitude: float = 0.2,
                        octaves: int = 6, seed: int = 0) -> Mesh:
    """Icosphere L``level`` of radius ``radius`` (metres), vertices displaced to
    ``radius * (1 + amplitude * fbm(v_hat))``; CCW outward; FP32 vertices."""
    v, f = icosphere(level)
    r = radius * (1.0 + amplitude * fbm(v, octaves=octaves, seed=seed))
    verts = (v * r[:, None]).astype(np.float32)
    return Mesh(verts, f.astype(np.uint32), "icosphere_L%d" % level)


def look_at_rotation(position, target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0)):
    """Rotation R for a camera at ``position`` whose -z boresight points at
    ``target``.  Rays are rotated by R^T (src/crt/cameras/simple_camera.hpp:71-76),
    so the rows of R are the camera axes in world coordinates."""
    p = np.asarray(position, dtype=np.float64)
    z = p - np.asarray(target, dtype=np.float64)
    z /= np.linalg.norm(z)
    up = np.asarray(up, dtype=np.float64)
    if abs(np.dot(up, z)) > 0.99:
        up = np.array([1.0, 0.0, 0.0])
    y = up - np.dot(up, z) * z
    y /= np.linalg.norm(y)
    x = np.cross(y, z)
    return np.stack([x, y, z])


def random_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def small_body_camera(mesh: Mesh, resolution=2048, seed=1, range_km=(50.0, 200.0)):
    """FP64 camera pose at a uniform-random range and direction looking at the
    origin; FOV = 2 atan(1.25 R_max / rho) so the body fills the frame."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(*range_km) * 1000.0
    pos = random_unit(rng) * rho
    rmax = float(np.max(np.linalg.norm(mesh.vertices.astype(np.float64), axis=1)))
    f = 1.0
    s = 2.0 * f * 1.25 * rmax / rho
    return {"focal_length": f, "resolution": [resolution, resolution], "sensor_size": [s, s],
            "z_positive": False, "position": pos, "rotation": look_at_rotation(pos)}


def camera_rays(cam: dict):
    """FP64 rays of every pixel (row-major, integer pixel coordinates) of a
    pinhole camera dict, as SimpleCamera::pixel_to_ray forms them
    (src/crt/cameras/simple_camera.hpp:61-81): normalised camera-frame
    direction ((u - cx) / sx, (cy - v) / sy, +-f), rotated by R^T.
    Returns (origins, dirs), each (H*W, 3)."""
    w, h = int(np.floor(cam["resolution"][0])), int(np.floor(cam["resolution"][1]))
    sx = cam["resolution"][0] / cam["sensor_size"][0]
    sy = cam["resolution"][1] / cam["sensor_size"][1]
    f = cam["focal_length"] if cam.get("z_positive", False) else -cam["focal_length"]
    v, u = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64), indexing="ij")
    dc = np.stack([(u.ravel() - cam["resolution"][0] / 2.0) / sx, (cam["resolution"][1] / 2.0 - v.ravel()) / sy,
                   np.full(w * h, float(f))], axis=1)
    dc /= np.linalg.norm(dc, axis=1, keepdims=True)
    R = np.asarray(cam["rotation"], dtype=np.float64).reshape(3, 3)
    d = dc @ R   # row k = R^T dc_k
    o = np.broadcast_to(np.asarray(cam["position"], dtype=np.float64), d.shape).copy()
    return o, d


def sun_direction(seed=2):
    """Unit vector from the body towards the sun (uniform on the sphere)."""
    return random_unit(np.random.default_rng(seed))


def geodesic_sphere(k: int, radius: float = 2000.0, amplitude: float = 0.2, octaves: int = 6,
                    seed: int = 0) -> Mesh:
    """Geodesic sphere of frequency k: every icosahedron face split into k^2
    triangles (20 k^2 in total), vertices projected to the sphere and displaced
    like displaced_icosphere.  Points on shared edges are computed from the
    same integer barycentric weights summed in global-vertex order, so they are
    bit-identical from both faces and deduplicate exactly (no cracks)."""
    v0, f0 = _icosahedron()
    ii, jj = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="ij")
    m = (ii + jj) <= k
    gi, gj = ii[m], jj[m]                     # grid points of one face
    lid = -np.ones((k + 1, k + 1), dtype=np.int64)
    lid[gi, gj] = np.arange(gi.size)
    ui, uj = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    um = (ui + uj) <= k - 1
    ui, uj = ui[um], uj[um]
    up = np.stack([lid[ui, uj], lid[ui + 1, uj], lid[ui, uj + 1]], 1)
    dm = (ui + uj) < k - 1
    di, dj = ui[dm], uj[dm]
    down = np.stack([lid[di + 1, dj], lid[di + 1, dj + 1], lid[di, dj + 1]], 1)
    local_tris = np.concatenate([up, down])
    pts, tris = [], []
    for fi in range(f0.shape[0]):
        a, b, c = f0[fi]
        w = np.stack([k - gi - gj, gi, gj], 1).astype(np.float64)   # weights of a, b, c
        order = np.argsort([a, b, c])
        verts = np.array([a, b, c])[order]
        ww = w[:, order]
        p = (ww[:, 0:1] * v0[verts[0]] + ww[:, 1:2] * v0[verts[1]]) + ww[:, 2:3] * v0[verts[2]]
        pts.append(p / k)
        tris.append(local_tris + fi * gi.size)
    P = np.concatenate(pts)
    T = np.concatenate(tris)
    uniq, inv = np.unique(P, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    T = inv[T]
    n = uniq / np.linalg.norm(uniq, axis=1, keepdims=True)
    r = radius * (1.0 + amplitude * fbm(n, octaves=octaves, seed=seed)) if amplitude else np.full(n.shape[0], radius)
    return Mesh((n * r[:, None]).astype(np.float32), T.astype(np.uint32), "geodesic_k%d" % k)
