"""Scene blobs: a built scene saved with crt_scene_serialize and loaded with
crt_scene_deserialize traces exactly like the original, keeps its frame, and
damaged blobs (checksum, truncation, magic, tree structure) are rejected
before anything is traced."""
import ctypes as C

import numpy as np
import pytest

import oracle
from crt import capi, scenes

pytestmark = pytest.mark.gpu


class _Header(C.Structure):
    # mirror of the blob header (csrc/scene_io.hip)
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("header_bytes", C.c_uint32),
                ("n_tris", C.c_uint32), ("n_vertices", C.c_uint32), ("n_instances", C.c_uint32),
                ("has_vnormals", C.c_uint32), ("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32),
                ("max_depth", C.c_uint32), ("pad_", C.c_uint32), ("has_mats", C.c_uint32),
                ("has_vuv", C.c_uint32), ("n_textures", C.c_uint32), ("pad2_", C.c_uint32),
                ("n_texels", C.c_uint64), ("origin", C.c_double * 3),
                ("center", C.c_double * 3), ("radius", C.c_double), ("bvh_bounds", C.c_float * 6),
                ("info", capi.BvhInfo), ("section_bytes", C.c_uint64 * 13), ("checksum", C.c_uint64)]


def _pad16(b):
    return (b + 15) & ~15


def _fnv(data, h=1469598103934665603):
    for x in data:
        h ^= int(x)
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _sections(blob):
    h = _Header.from_buffer_copy(blob[:C.sizeof(_Header)])
    off = _pad16(C.sizeof(_Header))
    out = []
    for b in h.section_bytes:
        out.append((off, b))
        off += _pad16(b)
    return h, out


@pytest.fixture(scope="module")
def cornell():
    sc = scenes.cornell_box(0)
    v, i, inst = sc.merged()
    vn = oracle.vertex_normals(v, i)
    s = capi.Scene(v, i, inst, vn, colors=np.array(sc.colors), smooth=np.array(sc.smooth, dtype=np.uint8))
    return s


def test_roundtrip_traces_identically(cornell, tmp_path):
    s = cornell
    path = tmp_path / "cornell.crtscene"
    s.save(path)
    r = capi.Scene.load(path)
    a, b = s.info(), r.info()
    for k in ("n_tris", "n_wide_nodes", "n_leaves", "max_depth", "bounds", "sphere", "stack_capacity"):
        assert a[k] == b[k], k
    cam = scenes.cornell_camera()
    pa, pb = s.camera_pass(cam), r.camera_pass(cam)
    for k in pa:
        if k != "timing":
            assert np.array_equal(pa[k], pb[k]), k
    rng = np.random.default_rng(0)
    o = rng.uniform(0.1, 0.9, size=(5000, 3))
    d = rng.normal(size=(5000, 3))
    ta = s.trace(o, d)
    tb = r.trace(o, d)
    for x, y in zip(ta, tb):
        assert np.array_equal(x, y)
    light = scenes.cornell_light()
    ia = s.render(cam, [light], 2, 2, 0.0, 2, seed=3)[0]
    ib = r.render(cam, [light], 2, 2, 0.0, 2, seed=3)[0]
    assert np.array_equal(ia, ib)


def test_roundtrip_keeps_far_frame():
    v0, f0 = scenes.icosphere(2)
    shift = np.array([3.0e6, 1.0e6, -2.0e6])
    v = (np.asarray(v0) * 50.0 + shift).astype(np.float32)
    s = capi.Scene(v, np.asarray(f0, dtype=np.uint32))
    r = capi.Scene.from_bytes(s.to_bytes())
    rng = np.random.default_rng(1)
    o = shift + rng.normal(size=(2000, 3)) * 400.0
    d = shift + rng.normal(size=(2000, 3)) * 20.0 - o
    ta, tb = s.trace(o, d), r.trace(o, d)
    assert np.array_equal(ta[1], tb[1]) and np.array_equal(ta[0], tb[0])
    assert (ta[1] >= 0).sum() > 500
    assert np.allclose(r.info()["sphere"][:3], s.info()["sphere"][:3])


def _expect_invalid(blob):
    # CRT_ERR_INVALID surfaces as ValueError
    with pytest.raises(ValueError) as e:
        capi.Scene.from_bytes(bytes(blob))
    return str(e.value)


def test_damaged_blobs_are_rejected(cornell):
    blob = bytearray(cornell.to_bytes())
    h, secs = _sections(blob)
    # a flipped byte in the triangle section
    bad = bytearray(blob)
    off, n = secs[8]
    bad[off + n // 2] ^= 0x40
    assert "checksum" in _expect_invalid(bad)
    # truncation and magic
    assert "truncated" in _expect_invalid(blob[:len(blob) // 2])
    bad = bytearray(blob)
    bad[0] ^= 0xFF
    assert "magic" in _expect_invalid(bad)
    # a child reference pointing back at the root, in both node layouts and
    # with the checksum recomputed: caught by the structure checks
    bad = bytearray(blob)
    o256, _ = secs[6]
    o96, _ = secs[7]
    refs = np.frombuffer(bytes(bad[o256:o256 + 256]), dtype=np.uint32)[3::8]
    k = int(np.where((refs != 0xFFFFFFFF) & (refs < 0x80000000))[0][0]) if np.any(refs < 0x80000000) else 0
    zero = np.uint32(0).tobytes()
    bad[o256 + 32 * k + 12:o256 + 32 * k + 16] = zero
    bad[o96 + 64 + 4 * k:o96 + 64 + 4 * k + 4] = zero
    hs = _Header.from_buffer_copy(bytes(bad[:C.sizeof(_Header)]))
    data = b"".join(bytes(bad[o:o + n]) for o, n in secs)
    hs.checksum = _fnv(np.frombuffer(data, dtype=np.uint8))
    bad[:C.sizeof(_Header)] = bytes(hs)
    msg = _expect_invalid(bad)
    assert "scene blob" in msg and "checksum" not in msg
