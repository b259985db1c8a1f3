"""Closed-form material scenes (test helper), shared by the oracle (CPU) and
GPU tests.  Each returns a dict: vertices, indices, tri_inst, colors,
materials (one dict per instance, crt.capi.make_material fields), uv,
textures, camera, lights, and what the reference's formulas fix
(src/crt/materials/material.hpp:50-60, 90-184;
src/crt/path_tracing/unidirectional.hpp:24-96)."""
import numpy as np

TWO_PI_F = np.float32(2 * np.pi)


def _quad(x0, x1, y0, y1, z, n_sign):
    """Two triangles of the rectangle [x0,x1] x [y0,y1] at height z whose
    reference normal (p0 - p1) x (p2 - p0) has the sign n_sign along z."""
    v = np.array([[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]], dtype=np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    n = np.cross(v[f[0, 0]] - v[f[0, 1]], v[f[0, 2]] - v[f[0, 0]])
    if np.sign(n[2]) != n_sign:
        f = f[:, ::-1].copy()
    return v, f


def _down_camera(res=64, height=3.0, half=0.8):
    # looks along -z (z_positive False, R = I) from above the plane z = 0
    return {"focal_length": 1.0, "resolution": [res, res], "sensor_size": [2 * half / height] * 2,
            "z_positive": False, "position": np.array([0.0, 0.0, height]), "rotation": np.eye(3)}


def lit_plane(material, color=(0.5, 0.25, 0.75), res=64):
    """One quad in z = 0 facing the camera and a directional light straight
    above (irradiance 0.8); the plane's reference normal points into it
    (-z), so -L.n = 1 for the light and every pixel is lit and unoccluded."""
    v, f = _quad(-1, 1, -1, 1, 0.0, -1)
    return {"vertices": v, "indices": f, "tri_inst": np.ones(2, np.uint32), "colors": np.array([color]),
            "materials": [dict(material, color=color)], "uv": None, "textures": [],
            "camera": _down_camera(res), "lights": [{"type": "directional", "intensity": 0.8,
                                                     "direction": np.array([0.0, 0.0, 1.0])}]}


def checker_plane(res=96, n=4):
    """A textured quad with UVs (0,0)..(1,1) over x, y in [-1, 1] and an
    n x n black / white checker: texel (tx, ty) of the hit follows
    tx = clamp(floor(u n + 0.5)), ty = clamp(floor(v n + 0.5))."""
    v, f = _quad(-1, 1, -1, 1, 0.0, -1)
    uv = np.stack([(v[:, 0] + 1) / 2, (v[:, 1] + 1) / 2], axis=1).astype(np.float32)
    tx, ty = np.meshgrid(np.arange(n), np.arange(n))
    tex = np.zeros((n, n, 4), np.uint8)
    tex[..., 3] = 255
    tex[(tx + ty) % 2 == 0, :3] = 255
    sc = lit_plane({"type": "textured_lambertian", "texture": 0}, res=res)
    sc.update(uv=uv, textures=[tex])
    return sc


def checker_expected(cam, n=4):
    """Per pixel: 1 (white texel), 0 (black), -1 (within 1e-3 texel of a texel
    boundary, or off the plane) for the checker_plane camera rays."""
    import oracle
    o, d = oracle.camera_rays(cam)
    t = -o[:, 2] / d[:, 2]
    p = o + t[:, None] * d
    out = -np.ones(o.shape[0])
    on = (np.abs(p[:, 0]) < 1) & (np.abs(p[:, 1]) < 1)
    u, w = (p[:, 0] + 1) / 2, (p[:, 1] + 1) / 2
    fx, fy = u * n + 0.5, w * n + 0.5
    tx = np.clip(np.floor(fx), 0, n - 1)
    ty = np.clip(np.floor(fy), 0, n - 1)
    near = (np.abs(fx - np.round(fx)) < 1e-3) | (np.abs(fy - np.round(fy)) < 1e-3)
    ok = on & ~near
    out[ok] = ((tx[ok] + ty[ok]) % 2 == 0).astype(float)
    r = int(cam["resolution"][0])
    return out.reshape(r, r)


def mirror_room(res=64):
    """A mirror floor (z = 0) under a ceiling at z = 2 made of a red half
    (y > 0) and a green half (y < 0), lit by a point light at z = 1; the
    camera at (0, -3.5, 1.2) looks down at the mirror.  Reflected camera rays
    (d - 2 n (d.n) = (dx, dy, -dz)) reach the ceiling at y = m_y + dy t,
    t = 2 / -dz, so each pixel is red or green by the sign of that y."""
    vf, ff = _quad(-6, 6, -6, 6, 0.0, -1)          # reference normal into the floor
    vr, fr = _quad(-6, 6, 0, 6, 2.0, +1)           # ceiling halves: normal into the ceiling
    vg, fg = _quad(-6, 6, -6, 0, 2.0, +1)
    v = np.concatenate([vf, vr, vg]).astype(np.float32)
    f = np.concatenate([ff, fr + 4, fg + 8]).astype(np.uint32)
    inst = np.array([1, 1, 2, 2, 3, 3], np.uint32)
    colors = np.array([(1, 1, 1), (0.9, 0.0, 0.0), (0.0, 0.9, 0.0)])
    mats = [{"type": "mirror"}, {"type": "lambertian", "color": colors[1]},
            {"type": "lambertian", "color": colors[2]}]
    pos = np.array([0.0, -3.5, 1.2])
    look = np.array([0.0, 0.5, 0.0]) - pos
    look /= np.linalg.norm(look)
    # rows of R: camera x, y, z axes in the world; rays go along -z_cam = look
    zc = -look
    xc = np.cross([0.0, 0.0, 1.0], zc)
    xc /= np.linalg.norm(xc)
    yc = np.cross(zc, xc)
    cam = {"focal_length": 1.0, "resolution": [res, res], "sensor_size": [0.9, 0.9], "z_positive": False,
           "position": pos, "rotation": np.stack([xc, yc, zc])}
    return {"vertices": v, "indices": f, "tri_inst": inst, "colors": colors, "materials": mats, "uv": None,
            "textures": [], "camera": cam,
            "lights": [{"type": "point", "intensity": 2.0, "position": np.array([0.3, 0.2, 1.0])}]}


def mirror_expected(cam):
    """+1 red / -1 green per pixel from the analytic reflection, 0 where the
    reflected ray lands within 1e-3 of y = 0 or misses the mirror / ceiling."""
    import oracle
    o, d = oracle.camera_rays(cam)
    out = np.zeros(o.shape[0])
    down = d[:, 2] < 0
    t0 = np.where(down, -o[:, 2] / np.where(down, d[:, 2], -1.0), 0.0)
    m = o + t0[:, None] * d
    ok = down & (np.abs(m[:, 0]) < 6) & (np.abs(m[:, 1]) < 6)
    t1 = 2.0 / np.where(ok, -d[:, 2], 1.0)
    cy = m[:, 1] + d[:, 1] * t1
    cx = m[:, 0] + d[:, 0] * t1
    ok &= (np.abs(cx) < 6) & (np.abs(cy) < 6) & (np.abs(cy) > 1e-3)
    out[ok] = np.sign(cy[ok])
    r = int(cam["resolution"][0])
    return out.reshape(r, r)


def mixed_cornell():
    """The Cornell box with a mirror red wall, a checker-textured green wall
    (UV = (z, y) on x = 1), a constant Blinn-Phong box and a textured
    Blinn-Phong box with a specular map: every material and map at once."""
    from crt import scenes
    sc = scenes.cornell_box(0)
    v, i, inst = sc.merged()
    rng = np.random.default_rng(3)
    uv = np.zeros((v.shape[0], 2), np.float32)
    base = 0
    for k, m in enumerate(sc.meshes):
        nv = m.vertices.shape[0]
        if k == 2:   # green wall, x = 1
            uv[base:base + nv] = np.stack([m.vertices[:, 2], m.vertices[:, 1]], axis=1)
        elif k == 4:
            uv[base:base + nv] = rng.uniform(0, 1, size=(nv, 2))
        base += nv
    tx, ty = np.meshgrid(np.arange(8), np.arange(8))
    checker = np.zeros((8, 8, 4), np.uint8)
    checker[..., 3] = 255
    checker[..., 1] = np.where((tx + ty) % 2 == 0, 220, 40)
    color_map = rng.integers(0, 256, size=(16, 16, 4), dtype=np.uint8)
    spec_map = rng.integers(0, 256, size=(16, 16, 4), dtype=np.uint8)
    cols = np.array(sc.colors)
    mats = [{"type": "lambertian", "color": cols[0]},
            {"type": "mirror"},
            {"type": "textured_lambertian", "texture": 0},
            {"type": "blinn_phong", "color": (0.8, 0.6, 0.3), "coeffs": (0.1, 0.7, 0.5), "alpha": 24.0},
            {"type": "blinn_phong", "texture": 1, "spec_texture": 2, "alpha": 16.0},
            {"type": "lambertian", "color": cols[5]}]
    return {"vertices": v, "indices": i, "tri_inst": inst, "colors": cols, "materials": mats, "uv": uv,
            "textures": [checker, color_map, spec_map], "camera": scenes.cornell_camera(),
            "lights": [dict(scenes.cornell_light(), type="area")]}


def build(sc, gpu):
    """(scene object with materials set) for the GPU (crt.capi) or the oracle."""
    if gpu:
        from crt import capi
        s = capi.Scene(sc["vertices"], sc["indices"], sc["tri_inst"], colors=sc["colors"])
    else:
        import oracle
        s = oracle.Scene(sc["vertices"], sc["indices"], sc["tri_inst"], colors=sc["colors"])
    s.set_materials(sc["materials"], sc["uv"], sc["textures"])
    return s
