"""Batch LiDAR pose ingestion past the reference's 100 000-pose cap
(src/crt/lidars/lidar.hpp:11): through the drop-in Python API with a
rotated, translated group, and from device-resident pose buffers, against
the FP64 oracle's lidar_batch (src/crt/do_lidar.hpp:80-155 semantics)."""
import numpy as np
import pytest

import oracle
from crt import capi, scenes

pytestmark = pytest.mark.gpu


def _poses(n, seed, center, r_lo, r_hi):
    """n poses on shells around `center`, boresight (-z of R^T) aimed at a
    point within 800 m of it; rotations (3, 3, n)."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    pos = center + u * rng.uniform(r_lo, r_hi, size=(n, 1))
    tgt = center + rng.normal(size=(n, 3)) * 800.0
    fwd = tgt - pos
    fwd /= np.linalg.norm(fwd, axis=1, keepdims=True)
    # R^T (0, 0, -1) = fwd -> third row of R = -fwd; complete a right-handed frame
    z = -fwd
    a = np.where(np.abs(z[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    x = np.cross(a, z)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(z, x)
    R = np.stack([x, y, z], axis=1)          # (n, 3, 3), rows x, y, z
    return pos, np.moveaxis(R, 0, 2)         # (n, 3), (3, 3, n)


def _rays(rel_p, rel_r):
    # boresight of SimpleLidar (z_positive False): R^T (0, 0, -1) = -(third row of R)
    return rel_p, -rel_r[2, :, :].T


def test_batch_simulate_lidar_past_cap_matches_oracle(tmp_path):
    from crt import BodyFixedEntity, BodyFixedGroup
    from crt.lidars import SimpleLidar
    import _crt
    v0, f0 = scenes.icosphere(5)
    mesh = scenes.Mesh(np.asarray(v0, dtype=np.float64) * 1000.0, np.asarray(f0, dtype=np.uint32))
    path = str(tmp_path / "body.obj")
    scenes.write_obj(path, mesh)
    ang = 0.7
    G = np.array([[np.cos(ang), -np.sin(ang), 0.0], [np.sin(ang), np.cos(ang), 0.0], [0.0, 0.0, 1.0]]) @ \
        np.array([[1.0, 0.0, 0.0], [0.0, np.cos(0.3), -np.sin(0.3)], [0.0, np.sin(0.3), np.cos(0.3)]])
    gpos = np.array([3.0e6, -1.2e6, 4.0e5])
    group = BodyFixedGroup(BodyFixedEntity(path), position=gpos, rotation=G)
    n = 150_000
    # world poses around the body's world position (group position)
    pos, rot = _poses(n, 7, gpos, 2.0e3, 2.0e4)
    lidar = SimpleLidar()
    lidar.batch_set_pose(pos, rot)
    ranges = np.asarray(group.batch_simulate_lidar(lidar)).ravel()
    assert ranges.shape == (n,)
    # the body-frame poses the library traced, and the same transform pose by pose
    rel_p, rel_r = group.transform_to_body_batch(pos, rot)
    for k in (0, 12345, n - 1):
        p1, r1 = group.transform_to_body(pos[k], rot[:, :, k])
        assert np.array_equal(p1, rel_p[k]) and np.array_equal(r1, rel_r[:, :, k])
    vp, fp, _ = _crt.load_obj(path)
    cpu = oracle.Scene(vp, fp)
    ref = cpu.lidar_batch(rel_p, np.moveaxis(rel_r, 2, 0).reshape(-1, 9))
    both = (ranges > 0) & (ref > 0)
    assert both.sum() > n // 4
    rel = np.abs(ranges[both] - ref[both]) / ref[both]
    bad = np.where((ranges > 0) != (ref > 0))[0]
    fragile = np.zeros(0, bool)
    if bad.size:
        o, d = _rays(rel_p[bad], rel_r[:, :, bad])
        fragile = cpu.closest_fragile(o, d, 2.0 ** -20 * 2000.0)
    print("batch lidar: %d poses, %d hits, hit-status disagreements %d (%d near an edge), max rel range err %.2e"
          % (n, both.sum(), bad.size, int(fragile.sum()), rel.max()))
    assert int((~fragile).sum()) == 0
    assert rel.max() < 1e-9


def test_device_pose_buffers_match_host_and_oracle():
    # 1.2M poses resident on the device: crt_lidar_batch with device inputs and
    # output gives the host path's ranges bit for bit; a sample matches the oracle
    m = scenes.displaced_icosphere(6, 2000.0, 0.2, 6, 0)
    sc = capi.Scene(m.vertices, m.indices)
    n = 1_200_000
    pos, rot = _poses(n, 11, np.zeros(3), 5.0e3, 5.0e4)
    rot9 = np.ascontiguousarray(np.moveaxis(rot, 2, 0).reshape(-1, 9))
    host = sc.lidar_batch(pos, rot9)
    dp = capi.DeviceBuffer(n * 24).from_host(np.ascontiguousarray(pos))
    dr = capi.DeviceBuffer(n * 72).from_host(rot9)
    do = capi.DeviceBuffer(n * 8)
    st = capi.Stream()
    sc.lidar_batch_device(n, dp.ptr, dr.ptr, do.ptr, stream=st.handle)
    st.synchronize()
    dev = do.to_host(np.empty(n))
    assert np.array_equal(dev, host)
    assert (host > 0).sum() > n // 4
    sel = np.random.default_rng(3).choice(n, 40_000, replace=False)
    cpu = oracle.Scene(m.vertices, m.indices)
    ref = cpu.lidar_batch(pos[sel], rot9[sel])
    both = (host[sel] > 0) & (ref > 0)
    bad = np.where((host[sel] > 0) != (ref > 0))[0]
    fragile = np.zeros(0, bool)
    if bad.size:
        o, d = _rays(pos[sel][bad], rot[:, :, sel][:, :, bad])
        fragile = cpu.closest_fragile(o, d, 2.0 ** -20 * 4800.0)
    rel = np.abs(host[sel][both] - ref[both]) / ref[both]
    print("device poses: %d, sample %d, disagreements %d (%d near an edge), max rel err %.2e"
          % (n, sel.size, bad.size, int(fragile.sum()), rel.max()))
    assert int((~fragile).sum()) == 0
    assert rel.max() < 1e-9
