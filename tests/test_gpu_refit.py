"""Refit for moving geometry (crt_scene_refit; (f) rank 1 of the dynamic
path: the reference rebuilds SweepSAH + PRO per call,
src/crt/rendering_dynamic/render.hpp:19-62, and ships a refitter,
lib/bvh/hierarchy_refitter.hpp:11-43).

* unmoved vertices refit to the built nodes bit for bit;
* a moved entity traces like the FP64 oracle built from the moved vertices
  (ids equal except ties / edge grazes / decisions within FP32 resolution,
  depth within 1e-5);
* moving bodies apart past the SAH growth limit rebuilds instead;
* the drop-in dynamic API refits its per-call scene when only poses change.
"""
import numpy as np
import pytest

import oracle
from crt import capi, scenes
from render_parity import fp32_eps

pytestmark = pytest.mark.gpu


def _sections(blob):
    import test_gpu_scene_blob as tb
    import ctypes as C
    h = tb._Header.from_buffer_copy(blob[:C.sizeof(tb._Header)])
    off = tb._pad16(C.sizeof(tb._Header))
    out = []
    for b in h.section_bytes:
        out.append(blob[off:off + b])
        off += tb._pad16(b)
    return out


def _two_bodies(shift=np.array([5000.0, 0.0, 0.0])):
    a = scenes.displaced_icosphere(6, 2000.0, 0.2, 6, 0)
    b = scenes.displaced_icosphere(5, 800.0, 0.2, 6, 1)
    vb = b.vertices.astype(np.float64) + shift
    v = np.concatenate([a.vertices.astype(np.float64), vb])
    f = np.concatenate([a.indices, b.indices + a.vertices.shape[0]]).astype(np.uint32)
    inst = np.concatenate([np.full(a.n_tris, 1, np.uint32), np.full(b.n_tris, 2, np.uint32)])
    return v, f, inst, a.vertices.shape[0]


def _compare(gpu, v64, f, inst, cam):
    cpu = oracle.Scene(v64, f, inst, f64=True)
    g = gpu.camera_pass(cam)
    c = cpu.camera_pass(cam)
    o, d = oracle.camera_rays(cam)
    mism = (g["tri"] != c["tri"]).ravel()
    expl = np.ones(0, bool)
    if mism.any():
        expl = cpu.tie_flags(o[mism], d[mism]) | cpu.edge_flags(o[mism], d[mism], g["tri"].ravel()[mism]) | \
            cpu.closest_fragile(o[mism], d[mism], fp32_eps(gpu.info()))
    hit = (~mism) & (c["tri"].ravel() >= 0)
    rel = np.abs(g["depth"].ravel()[hit] - c["depth"].ravel()[hit]) / c["depth"].ravel()[hit]
    print("pixels %d, hits %d, id mismatches %d (explained %d), depth max rel %.2e"
          % (mism.size, hit.sum(), mism.sum(), expl.sum(), rel.max()))
    assert expl.all()
    assert rel.max() < 1e-5
    return hit.sum()


def _rot(ax, ang):
    c, s = np.cos(ang), np.sin(ang)
    i, j = [(1, 2), (2, 0), (0, 1)][ax]
    R = np.eye(3)
    R[i, i] = c; R[j, j] = c; R[i, j] = -s; R[j, i] = s
    return R


def test_refit_unmoved_reproduces_build_bit_for_bit():
    v, f, inst, _ = _two_bodies()
    s = capi.Scene(v, f, inst, f64=True)
    before = _sections(s.to_bytes())
    assert s.refit(v) is False
    after = _sections(s.to_bytes())
    for k in (0, 6, 7, 8):   # vertices, both node layouts, leaf-ordered triangles
        assert before[k] == after[k], k
    assert s.info()["refit_ms"] > 0


def test_refit_moved_entity_matches_oracle():
    v, f, inst, na = _two_bodies()
    s = capi.Scene(v, f, inst, f64=True)
    info0 = s.info()
    # move the small body: rotate about its own centre and translate
    c0 = v[na:].mean(axis=0)
    v2 = v.copy()
    v2[na:] = (v[na:] - c0) @ _rot(2, 0.7).T @ _rot(0, 0.3).T + c0 + np.array([-800.0, 1500.0, 600.0])
    assert s.refit(v2) is False
    info = s.info()
    print("refit %.3f ms (build %.3f ms), SAH %.3f -> %.3f" % (info["refit_ms"], info0["build_ms"],
                                                                 info0["sah_cost"], info["sah_cost"]))
    for seed in (3, 4):
        cam = scenes.small_body_camera(scenes.Mesh(v2.astype(np.float32), f), resolution=256, seed=seed,
                                       range_km=(20.0, 40.0))
        assert _compare(s, v2, f, inst, cam) > 3000
    # the same scene built from scratch at the new pose traces the same rays alike
    fresh = capi.Scene(v2, f, inst, f64=True)
    cam = scenes.small_body_camera(scenes.Mesh(v2.astype(np.float32), f), resolution=128, seed=5)
    a, b = s.camera_pass(cam), fresh.camera_pass(cam)
    same = a["tri"] == b["tri"]
    assert same.mean() > 0.999
    assert np.allclose(a["depth"][same], b["depth"][same], rtol=1e-12, atol=0)


def test_refit_rebuilds_when_sah_grows():
    # the even and the odd triangles of one body as two entities: every node
    # of the tree mixes them, so moving one entity away stretches every node
    # across the gap and the refit's SAH grows past the limit: rebuilt.
    # (Bodies apart sit in separate subtrees, where a refit is as good as a
    # build: the moved-entity test above keeps its refit.)
    m = scenes.displaced_icosphere(6, 2000.0, 0.2, 6, 0)
    v0 = m.vertices.astype(np.float64)
    nv = v0.shape[0]
    v = np.concatenate([v0, v0])
    f = np.concatenate([m.indices[0::2], m.indices[1::2] + nv]).astype(np.uint32)
    inst = np.concatenate([np.full(len(m.indices[0::2]), 1, np.uint32), np.full(len(m.indices[1::2]), 2, np.uint32)])
    s = capi.Scene(v, f, inst, f64=True)
    sah0 = s.info()["sah_cost"]
    v2 = v.copy()
    v2[nv:] += np.array([6000.0, -4000.0, 3000.0])
    assert s.refit(v2) is True
    print("interleaved entities moved apart: rebuilt, SAH %.2f -> %.2f" % (sah0, s.info()["sah_cost"]))
    for part in (slice(0, nv), slice(nv, None)):
        c = v2[part].mean(axis=0)
        cam = scenes.small_body_camera(scenes.Mesh((v2[part] - c).astype(np.float32), f[:1]), resolution=128,
                                       seed=6, range_km=(10.0, 20.0))
        cam["position"] = cam["position"] + c
        assert _compare(s, v2, f, inst, cam) > 1000


def test_dynamic_api_refits_moved_entity(tmp_path):
    import _crt
    from crt import Entity, rendering
    from crt.cameras import SimpleCamera
    a = scenes.displaced_icosphere(5, 2000.0, 0.2, 6, 0)
    b = scenes.displaced_icosphere(4, 600.0, 0.2, 6, 1)
    pa, pb = str(tmp_path / "a.obj"), str(tmp_path / "b.obj")
    scenes.write_obj(pa, a)
    scenes.write_obj(pb, b)
    ea, eb = Entity(pa), Entity(pb, position=[4000.0, 0.0, 0.0])
    cam_pos = np.array([1500.0, -25000.0, 0.0])
    look = -cam_pos / np.linalg.norm(cam_pos)
    zc = -look
    xc = np.cross([0.0, 0.0, 1.0], zc)
    xc /= np.linalg.norm(xc)
    R = np.stack([xc, np.cross(zc, xc), zc])
    mk = lambda: SimpleCamera(1.0, [160, 160], [0.3, 0.3], False, position=cam_pos, rotation=R)
    st0 = _crt.dynamic_scene_stats()
    rendering.instance_pass(mk(), [ea, eb])
    eb.set_pose([3500.0, 800.0, 300.0], _rot(2, 0.5))
    ids = rendering.instance_pass(mk(), [ea, eb])
    xyz = rendering.intersection_pass(mk(), [ea, eb])
    st1 = _crt.dynamic_scene_stats()
    assert st1["refits"] - st0["refits"] == 2 and st1["builds"] - st0["builds"] <= 1
    # oracle from the moved FP64 vertices of the OBJ-parsed meshes
    va, fa, _ = _crt.load_obj(pa)
    vb, fb, _ = _crt.load_obj(pb)
    vb = vb @ _rot(2, 0.5).T + np.array([3500.0, 800.0, 300.0])
    v = np.concatenate([va, vb])
    f = np.concatenate([fa, fb + va.shape[0]]).astype(np.uint32)
    inst = np.concatenate([np.full(fa.shape[0], 1, np.uint32), np.full(fb.shape[0], 2, np.uint32)])
    cpu = oracle.Scene(v, f, inst, f64=True)
    cam = {"position": cam_pos, "rotation": R, "focal_length": 1.0, "resolution": [160, 160],
           "sensor_size": [0.3, 0.3], "z_positive": False}
    c = cpu.camera_pass(cam)
    same = ids == c["instance"]
    print("dynamic refit: %d pixels on body 2, instance agreement %.5f" % ((c["instance"] == 2).sum(), same.mean()))
    assert (c["instance"] == 2).sum() > 300
    assert same.mean() > 0.9995
    hit = same & (c["tri"] >= 0)
    err = np.linalg.norm(xyz[hit] - c["xyz"][hit], axis=1) / c["depth"][hit]
    assert err.max() < 1e-5
