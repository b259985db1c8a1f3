"""The drop-in `crt` API on the GPU, checked against the reference's own outputs.

BodyFixedGroup built from OBJ files (like a reference user would), passes
compared with tests/golden/cornell_passes.npz (written by the reference
tracer) and LiDAR with tests/golden/lidar_batch.npz.
"""
import os

import numpy as np
import pytest

import oracle
from crt import scenes

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def cornell_group(tmp_path_factory):
    from crt import BodyFixedEntity, BodyFixedGroup
    d = tmp_path_factory.mktemp("cornell")
    sc = scenes.cornell_box(0)
    ents = []
    for k, m in enumerate(sc.meshes):
        p = str(d / ("e%d.obj" % k))
        scenes.write_obj(p, m)
        ents.append(BodyFixedEntity(p, color=list(sc.colors[k])))
    return sc, BodyFixedGroup(ents), d


def _camera():
    from crt.cameras import SimpleCamera
    c = scenes.cornell_camera()
    return SimpleCamera(c["focal_length"], c["resolution"], c["sensor_size"], c["z_positive"],
                        position=c["position"], rotation=c["rotation"])


def test_passes_match_reference_golden(cornell_group):
    sc, group, _ = cornell_group
    g = np.load(os.path.join(GOLD, "cornell_passes.npz"))
    inst = group.instance_pass(_camera())
    xyz = group.intersection_pass(_camera())
    nrm = group.normal_pass(_camera())
    assert inst.dtype == np.uint32 and inst.shape == (256, 256)
    assert xyz.dtype == np.float64 and xyz.shape == (256, 256, 3)
    tri = group.triangle_pass(_camera())
    mism = tri != g["tri"]
    v, i, ids = sc.merged()
    cpu = oracle.Scene(v, i, ids)
    o, d = oracle.camera_rays(scenes.cornell_camera())
    m = mism.ravel()
    expl = cpu.tie_flags(o[m], d[m]) | cpu.edge_flags(o[m], d[m], g["tri"].ravel()[m]) | \
        cpu.edge_flags(o[m], d[m], tri.ravel()[m])
    print("API vs reference: %d id mismatches, %d explained as ties/edges" % (m.sum(), expl.sum()))
    assert np.all(expl)
    same = ~mism
    assert np.array_equal(inst[same], g["instance"][same])
    hit = same & (g["tri"] >= 0)
    depth = np.linalg.norm(g["xyz"][hit].astype(np.float64) - scenes.cornell_camera()["position"], axis=1)
    err = np.linalg.norm(xyz[hit] - g["xyz"][hit], axis=1) / depth
    assert err.max() < 1e-5
    assert np.max(np.abs(nrm[hit] - g["normal"][hit])) < 1e-5
    assert np.all(xyz[same & (g["tri"] < 0)] == 0) and np.all(inst[same & (g["tri"] < 0)] == 0)


def test_camera_pose_written_back_like_reference(cornell_group):
    sc, group, _ = cornell_group
    cam = _camera()
    group.set_pose([0.0, 0.0, 0.0], np.eye(3))
    group.instance_pass(cam)
    assert np.allclose(cam.position, scenes.cornell_camera()["position"])


def test_render_uint8_and_radiance(cornell_group):
    from crt.lights import AreaLight
    sc, group, _ = cornell_group
    L = scenes.cornell_light()
    light = AreaLight(L["intensity"], L["size"], position=L["position"], rotation=L["rotation"])
    cam = _camera()
    img = group.render(cam, light, 4, 4, 0.0, 3, seed=5)
    assert img.dtype == np.uint8 and img.shape == (256, 256, 4)
    assert np.all(img[..., 3] == 255)
    rad = group.render_radiance(_camera(), [light], 4, 4, 0.0, 3, seed=5)
    assert rad.dtype == np.float32 and rad.shape == (256, 256, 4)
    assert np.array_equal(img[..., :3], np.clip(rad[..., :3] * 256, 0, 255).astype(np.uint8))
    # same seed -> same image, different seed -> different noise
    assert np.array_equal(img, group.render(_camera(), light, 4, 4, 0.0, 3, seed=5))
    assert not np.array_equal(img, group.render(_camera(), light, 4, 4, 0.0, 3, seed=6))


def test_dynamic_api_matches_body_fixed(cornell_group):
    from crt import Entity
    from crt import rendering
    sc, group, d = cornell_group
    ents = [Entity(str(d / ("e%d.obj" % k)), color=list(sc.colors[k])) for k in range(len(sc.meshes))]
    a = rendering.instance_pass(_camera(), ents)
    b = group.instance_pass(_camera())
    assert np.array_equal(a, b)
    # moving an entity moves its hits (the dynamic API rebuilds the BVH per call)
    ents[3].set_position([0.0, 0.0, 0.2])
    c = rendering.instance_pass(_camera(), ents)
    assert not np.array_equal(a, c)


def test_lidar_api_matches_reference_golden(tmp_path):
    from crt import BodyFixedEntity, BodyFixedGroup
    from crt.lidars import SimpleLidar
    m = scenes.displaced_icosphere(8)
    p = str(tmp_path / "ico.obj")
    scenes.write_obj(p, m)
    group = BodyFixedGroup(BodyFixedEntity(p))
    g = np.load(os.path.join(GOLD, "lidar_batch.npz"))
    lid = SimpleLidar()
    lid.batch_set_pose(g["position"], np.transpose(g["rotation"], (1, 2, 0)))
    r = group.batch_simulate_lidar(lid)
    assert r.shape == (1, g["position"].shape[0])
    ref = g["range"]
    both = (r[0] > 0) & (ref > 0)
    assert ((r[0] > 0) != (ref > 0)).sum() <= 1
    assert np.max(np.abs(r[0][both] - ref[both]) / ref[both]) < 1e-5
    single = SimpleLidar(position=g["position"][0], rotation=g["rotation"][0])
    assert group.simulate_lidar(single) == pytest.approx(g["single_range"][0], rel=1e-5)


def test_unknown_light_type_raises(cornell_group):
    sc, group, _ = cornell_group
    with pytest.raises(TypeError):
        group._cpp.render(_camera()._cpp, [object()], 1, 1, 1.0, 1, None)


def test_rotated_translated_group_matches_reference_golden(cornell_group):
    """A group posed away from the identity (rotated about two axes and
    translated 1e5 m) seen by a camera whose world pose is the closed-form
    image of the golden camera pose, p_w = p_g + R_g^T q and R_w = R_c R_g:
    transform_to_body maps it back to (q, R_c), so the passes equal the
    reference's own identity-pose golden passes (crt/body_fixed.py:87-101)."""
    from crt.cameras import SimpleCamera
    sc, group, _ = cornell_group
    g = np.load(os.path.join(GOLD, "cornell_passes.npz"))
    c0 = scenes.cornell_camera()
    a, b = 0.7, -0.4
    Rg = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]]) @ \
        np.array([[1.0, 0.0, 0.0], [0.0, np.cos(b), -np.sin(b)], [0.0, np.sin(b), np.cos(b)]])
    pg = np.array([1.0e5, -3.0e4, 2.0e4])
    group.set_pose(pg, Rg)
    try:
        Rc = np.asarray(c0["rotation"], dtype=np.float64)
        pw = pg + Rg.T @ np.asarray(c0["position"], dtype=np.float64)
        def cam():
            return SimpleCamera(c0["focal_length"], c0["resolution"], c0["sensor_size"], c0["z_positive"],
                                position=pw, rotation=Rc @ Rg)
        c = cam()
        tri = group.triangle_pass(c)
        # the body-frame pose is written back into the caller's camera, like the reference
        assert np.allclose(c.position, c0["position"], atol=1e-9)
        assert np.allclose(c.rotation, Rc, atol=1e-14)
        xyz = group.intersection_pass(cam())
    finally:
        group.set_pose(np.zeros(3), np.eye(3))
    mism = (tri != g["tri"]).ravel()
    v, i, ids = sc.merged()
    cpu = oracle.Scene(v, i, ids)
    o, d = oracle.camera_rays(c0)
    expl = cpu.tie_flags(o[mism], d[mism]) | cpu.edge_flags(o[mism], d[mism], g["tri"].ravel()[mism]) | \
        cpu.edge_flags(o[mism], d[mism], tri.ravel()[mism])
    print("posed group vs reference: %d id mismatches, %d explained as ties/edges" % (mism.sum(), expl.sum()))
    assert np.all(expl)
    hit = (~mism.reshape(tri.shape)) & (g["tri"] >= 0)
    err = np.linalg.norm(xyz[hit] - g["xyz"][hit], axis=1)
    assert err.max() < 1e-5
