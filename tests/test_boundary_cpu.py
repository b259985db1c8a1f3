"""Boundary checks that need no GPU: the C-ABI library loads and exports every
symbol include/crt_hip.h declares, the _crt host module mirrors the reference
binding's names, and the OBJ loader follows the reference loader."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared():
    with open(os.path.join(ROOT, "include", "crt_hip.h")) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(crt_\w+)\s*\(", txt, re.M)))


def test_c_abi_exports_every_declared_symbol():
    import ctypes
    from crt import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = _declared()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_pybind_module_mirrors_reference_names():
    import _crt
    for cls in ["SimpleCamera", "SimpleLidar", "PointLight", "AreaLight", "Entity", "BodyFixedEntity",
                "BodyFixedGroup"]:
        assert hasattr(_crt, cls)
    for fn in ["render", "simulate_lidar", "intersection_pass", "instance_pass", "normal_pass"]:
        assert callable(getattr(_crt, fn))
    g = _crt.BodyFixedGroup
    for meth in ["render", "simulate_lidar", "batch_simulate_lidar", "intersection_pass", "instance_pass",
                 "normal_pass", "set_scale", "set_position", "set_rotation", "set_pose"]:
        assert hasattr(g, meth)


def test_pose_setters_and_validation():
    from crt.cameras import SimpleCamera
    from crt.lidars import SimpleLidar
    c = SimpleCamera(1.0, [8, 4], [1.0, 0.5], position=[1, 2, 3])
    assert np.array_equal(c._cpp.position, [1, 2, 3])
    R = np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 1.0]])
    c.set_rotation(R)
    assert np.array_equal(c._cpp.rotation, R)
    with pytest.raises(ValueError):
        c.set_rotation(np.ones((3, 3)))
    with pytest.raises(ValueError):
        c.set_position([1, 2])
    l = SimpleLidar()
    with pytest.raises(ValueError):
        l.batch_set_pose(np.zeros((3, 3)), np.zeros((3, 3, 2)))


def test_obj_loader_matches_reference_loader(tmp_path):
    """Vertex values (strtof), fan triangulation and the inward area-weighted
    vertex normals of lib/model_loaders/obj.hpp:57-117, checked against the
    reference loader's output kept in the Cornell fixture."""
    import _crt
    from crt import scenes
    sc = scenes.cornell_box(0)
    ref_vn = np.load(os.path.join(GOLD, "cornell_passes.npz"))["tri_vertex_normals"].reshape(-1, 3, 3)
    got = []
    for k, m in enumerate(sc.meshes):
        p = str(tmp_path / ("e%d.obj" % k))
        scenes.write_obj(p, m)
        v, f, vn = _crt.load_obj(p)
        assert np.array_equal(v.astype(np.float32), m.vertices)
        assert np.array_equal(f, m.indices)
        got.append(vn[f])
    assert np.allclose(np.concatenate(got), ref_vn, atol=1e-12)


def test_obj_loader_fan_negative_and_slashes(tmp_path):
    import _crt
    p = tmp_path / "poly.obj"
    p.write_text("# quad + pentagon\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0\n"
                 "f 1/1/1 2/2/2 3/3/3 4/4/4\nf -5//1 -4//1 -3//1 -1//1 -2//1\n\n")
    v, f, vn = _crt.load_obj(str(p))
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 2], [0, 2, 4], [0, 4, 3]]
    # CCW seen from +z: reference normal (p0-p1)x(p2-p0) points to -z
    assert np.allclose(vn[:3], [[0, 0, -1]] * 3)
    with pytest.raises(ValueError):
        _crt.load_obj(str(tmp_path / "missing.obj"))
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(ValueError):
        _crt.load_obj(str(bad))


def test_unsupported_geometry_type_rejected(tmp_path):
    from crt.entity import Entity
    p = tmp_path / "t.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(ValueError):
        Entity(str(p), geometry_type="ply")


def _rot(axis, ang):
    c, s = np.cos(ang), np.sin(ang)
    if axis == 0:
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == 1:
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_transform_to_body_closed_form():
    """BodyFixedGroup.transform_to_body (crt/body_fixed.py:87-101 of the
    reference): p' = scale * R_g (p - p_g), R' = (R_g R^T)^T, pinned by a
    hand-computed case and by round trips through the inverse map on rotated,
    translated and scaled group poses.  A transposed R_g, a dropped scale or
    a missing final transpose fails one of them."""
    from types import SimpleNamespace
    from crt.body_fixed import BodyFixedGroup
    f = BodyFixedGroup.transform_to_body
    # hand-computed: group turned 90 deg about z, at (10, 0, 0), scale 2
    g = SimpleNamespace(position=np.array([10.0, 0.0, 0.0]), rotation=_rot(2, np.pi / 2), scale=2.0)
    p, R = f(g, np.array([10.0, 5.0, 0.0]), np.eye(3))
    assert np.allclose(p, [-10.0, 0.0, 0.0], atol=1e-12)
    assert np.allclose(R, [[0, 1, 0], [-1, 0, 0], [0, 0, 1]], atol=1e-15)
    # a camera rotated 30 deg about x in the world, group turned about y
    g = SimpleNamespace(position=np.array([0.0, 0.0, 0.0]), rotation=_rot(1, 0.25), scale=1.0)
    Rw = _rot(0, np.pi / 6)
    _, R = f(g, np.zeros(3), Rw)
    assert np.allclose(R, Rw @ _rot(1, 0.25).T, atol=1e-15)
    # a body-frame ray direction is R_g times the world one: R'^T e = R_g R_w^T e
    e = np.array([0.0, 0.0, -1.0])
    assert np.allclose(R.T @ e, _rot(1, 0.25) @ (Rw.T @ e), atol=1e-15)
    # round trips: the world pose of a known body-frame pose (q, Rc) maps back
    rng = np.random.default_rng(0)
    for scale in (1.0, 2.5, 1e-3):
        for _ in range(20):
            Rg, _r = np.linalg.qr(rng.normal(size=(3, 3)))
            Rg = Rg * np.sign(np.linalg.det(Rg))
            Rc, _r = np.linalg.qr(rng.normal(size=(3, 3)))
            Rc = Rc * np.sign(np.linalg.det(Rc))
            pg = rng.normal(size=3) * 1e5
            q = rng.normal(size=3) * 1e3
            g = SimpleNamespace(position=pg, rotation=Rg, scale=scale)
            p, R = f(g, pg + Rg.T @ (q / scale), Rc @ Rg)
            assert np.allclose(p, q, rtol=1e-9, atol=1e-9 * np.linalg.norm(pg) * scale)
            assert np.allclose(R, Rc, atol=1e-14)
