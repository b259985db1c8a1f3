"""GPU parity of the camera passes and general ray queries against the FP64 oracle.

Tolerances (from the task statement for this path): triangle / instance ids
bit-exact except for rays the oracle classifies as exact ties or edge hits
(reported as a fraction); depth and hit points within 1e-5 relative.
"""
import numpy as np
import pytest

import oracle
from crt import capi, scenes
from render_parity import fp32_eps

pytestmark = pytest.mark.gpu

REL = 1e-5


def _cornell():
    sc = scenes.cornell_box(0)
    v, i, inst = sc.merged()
    return sc, v, i, inst


@pytest.fixture(scope="module")
def cornell():
    sc, v, i, inst = _cornell()
    vn = oracle.vertex_normals(v, i)
    gpu = capi.Scene(v, i, inst, vn, colors=np.array(sc.colors), smooth=np.array(sc.smooth, dtype=np.uint8))
    cpu = oracle.Scene(v, i, inst, vn, colors=np.array(sc.colors), smooth=np.array(sc.smooth, dtype=np.uint8))
    return gpu, cpu, v, i


@pytest.fixture(scope="module")
def icosphere():
    m = scenes.displaced_icosphere(8)
    gpu = capi.Scene(m.vertices, m.indices)
    cpu = oracle.Scene(m.vertices, m.indices)
    return gpu, cpu, m


def _compare_pass(gpu_out, cpu_out, cpu, cam, eps):
    # every id that differs must be an exact tie, an edge/vertex graze
    # (barycentric within 1e-5), or a closest-hit decision within eps of an
    # edge or of another triangle (FP32 resolution; crt_oracle_closest_fragile)
    o, d = oracle.camera_rays(cam)
    tri_g = gpu_out["tri"].ravel()
    tri_c = cpu_out["tri"].ravel()
    mism = tri_g != tri_c
    ties = np.zeros(0, bool)
    fr = np.zeros(0, bool)
    if mism.any():
        ties = cpu.tie_flags(o[mism], d[mism]) | cpu.edge_flags(o[mism], d[mism], tri_g[mism])
        fr = cpu.closest_fragile(o[mism], d[mism], eps)
    hard = int((~(ties | fr)).sum())
    n = tri_c.size
    print("pixels %d  mismatched ids %d  tie/edge %d  within eps %d  unexplained %d"
          % (n, mism.sum(), ties.sum(), (fr & ~ties).sum(), hard))
    assert hard == 0
    same = ~mism
    hit = same & (tri_c >= 0)
    dg = gpu_out["depth"].ravel()[hit]
    dc = cpu_out["depth"].ravel()[hit]
    rel = np.abs(dg - dc) / np.maximum(dc, 1e-30)
    print("depth max rel err %.3e" % (rel.max() if rel.size else 0))
    assert rel.size == 0 or rel.max() < REL
    xg = gpu_out["xyz"].reshape(-1, 3)[hit]
    xc = cpu_out["xyz"].reshape(-1, 3)[hit]
    err = np.linalg.norm(xg - xc, axis=1) / np.maximum(dc, 1e-30)
    assert err.size == 0 or err.max() < REL
    # misses are zeros, like the reference (src/crt/passes.hpp:60-63)
    miss = same & (tri_c < 0)
    assert np.all(gpu_out["xyz"].reshape(-1, 3)[miss] == 0.0)
    assert np.all(gpu_out["depth"].ravel()[miss] == 0.0)
    return mism.mean(), ties.mean() if ties.size else 0.0


def test_cornell_passes_match_oracle(cornell):
    gpu, cpu, v, i = cornell
    cam = scenes.cornell_camera()
    g = gpu.camera_pass(cam)
    c = cpu.camera_pass(cam)
    _compare_pass(g, c, cpu, cam, fp32_eps(gpu.info()))
    same = g["tri"] == c["tri"]
    assert np.array_equal(g["instance"][same], c["instance"][same])
    hit = same & (c["tri"] >= 0)
    assert np.max(np.abs(g["normal"][hit] - c["normal"][hit])) < 1e-6
    assert np.all(g["instance"][same & (c["tri"] < 0)] == 0)


def test_cornell_smooth_normals(cornell):
    sc, v, i, inst = _cornell()
    vn = oracle.vertex_normals(v, i)
    smooth = np.ones(len(sc.meshes), dtype=np.uint8)
    gpu = capi.Scene(v, i, inst, vn, colors=np.array(sc.colors), smooth=smooth)
    cpu = oracle.Scene(v, i, inst, vn, colors=np.array(sc.colors), smooth=smooth)
    cam = scenes.cornell_camera()
    g = gpu.camera_pass(cam)
    c = cpu.camera_pass(cam)
    hit = (g["tri"] == c["tri"]) & (c["tri"] >= 0)
    assert np.max(np.abs(g["normal"][hit] - c["normal"][hit])) < 1e-5


def test_icosphere_primary_subsample(icosphere):
    gpu, cpu, m = icosphere
    cam = scenes.small_body_camera(m, resolution=512, seed=1)
    g = gpu.camera_pass(cam)
    c = cpu.camera_pass(cam)
    frac, tie = _compare_pass(g, c, cpu, cam, fp32_eps(gpu.info()))
    print("icosphere 512^2: id mismatch fraction %.2e (tie/edge %.2e)" % (frac, tie))


def test_trace_rays_closest_and_any(icosphere):
    gpu, cpu, m = icosphere
    rng = np.random.default_rng(7)
    n = 20000
    # rays from a shell at 3-6 km towards random points near the body
    o = rng.normal(size=(n, 3))
    o *= (rng.uniform(3000, 6000, size=n) / np.linalg.norm(o, axis=1))[:, None]
    tgt = rng.normal(size=(n, 3)) * 800.0
    d = tgt - o
    d *= rng.uniform(0.5, 3.0, size=n)[:, None]     # non-unit directions: t is in units of |d|
    tg, trg, ug, vg = gpu.trace(o, d)
    tc, trc, uc, vc = cpu.trace(o, d)
    mism = trg != trc
    ties = (cpu.tie_flags(o[mism], d[mism]) | cpu.edge_flags(o[mism], d[mism], trg[mism])) \
        if mism.any() else np.zeros(0, bool)
    assert mism.sum() - ties.sum() <= 2
    hit = (~mism) & (trc >= 0)
    assert np.max(np.abs(tg[hit] - tc[hit]) / tc[hit]) < REL
    assert np.max(np.abs(ug[hit] - uc[hit])) < 1e-3 and np.max(np.abs(vg[hit] - vc[hit])) < 1e-3
    # any-hit with tmax just before / after the closest hit
    sel = np.where(hit)[0][:2000]
    tmax_short = tc[sel] * (1 - 1e-4)
    tmax_long = tc[sel] * (1 + 1e-4)
    _, a_short, _, _ = gpu.trace(o[sel], d[sel], tmax=tmax_short, any_hit=True)
    _, a_long, _, _ = gpu.trace(o[sel], d[sel], tmax=tmax_long, any_hit=True)
    _, c_short, _, _ = cpu.trace(o[sel], d[sel], tmax=tmax_short, any_hit=True)
    assert np.all(a_long >= 0)
    assert np.array_equal(a_short >= 0, c_short >= 0)


def test_sun_pass_shadows(icosphere):
    gpu, cpu, m = icosphere
    cam = scenes.small_body_camera(m, resolution=256, seed=1)
    sun = scenes.sun_direction(2)
    r = gpu.sun_pass(cam, sun)
    c = cpu.camera_pass(cam)
    hit = c["tri"] >= 0
    assert r["timing"]["shadow_rays"] == int((r["tri"] >= 0).sum())
    # oracle shadow rays from the oracle hit points, offset like the GPU
    p = c["xyz"].reshape(-1, 3)[hit.ravel()]
    n = c["normal"].reshape(-1, 3)[hit.ravel()]
    o = p - gpu.info()["self_offset"] * n
    dd = np.repeat(np.asarray(sun)[None, :], o.shape[0], axis=0)
    _, occ, _, _ = cpu.trace(o, dd, any_hit=True)
    lit_c = occ < 0
    lit_g = r["lit"].ravel()[hit.ravel()].astype(bool)
    same_id = r["tri"].ravel()[hit.ravel()] == c["tri"].ravel()[hit.ravel()]
    disagree = (lit_c != lit_g) & same_id
    # every disagreement is a terminator or an edge within FP32 resolution:
    # the shadow ray's occlusion flips between triangles grown and shrunk by eps
    eps = fp32_eps(gpu.info())
    fr = cpu.shadow_fragile(o[disagree], dd[disagree], eps)
    print("shadow rays %d, disagreements %d, all within eps (terminator / edge): %s; fragile rays overall %.4f"
          % (hit.sum(), disagree.sum(), bool(fr.all()), cpu.shadow_fragile(o, dd, eps).mean()))
    assert fr.all()
    assert disagree.mean() < 2e-3
    assert np.all(r["lit"][~hit] == 0)


def _hip():
    """hipMalloc / hipMemcpy / hipFree / hipDeviceSynchronize of the HIP
    runtime libcrt_hip.so is bound to (looked up through its handle)."""
    import ctypes as C
    L = capi.load()
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    return L


def test_sun_pass_device_outputs_enqueue_only(icosphere):
    # device outputs without timing: the call only enqueues work on the
    # stream; several passes back to back, one synchronisation, then the
    # buffers must hold exactly the host-output result
    import ctypes as C
    gpu, _, m = icosphere
    n = 192
    cam = scenes.small_body_camera(m, resolution=n, seed=1)
    sun = scenes.sun_direction(2)
    ref = gpu.sun_pass(cam, sun)
    L = _hip()
    bufs = {}
    for name, size in (("depth", 8), ("tri", 4), ("lit", 1)):
        p = C.c_void_p()
        assert L.hipMalloc(C.byref(p), n * n * size) == 0
        assert L.hipMemset(p, 0x5A, n * n * size) == 0
        bufs[name] = p
    try:
        outs = {k: v.value for k, v in bufs.items()}
        for _ in range(4):
            assert gpu.sun_pass(cam, sun, outs, device=True, stream=None, timing=False) is None
        assert L.hipDeviceSynchronize() == 0
        got = {"depth": np.zeros((n, n)), "tri": np.zeros((n, n), np.int32), "lit": np.zeros((n, n), np.uint8)}
        for k, a in got.items():
            assert L.hipMemcpy(a.ctypes.data, bufs[k], a.nbytes, 2) == 0   # device to host
            assert np.array_equal(a, ref[k]), k
    finally:
        for p in bufs.values():
            L.hipFree(p)


def test_sun_passes_on_two_streams(icosphere):
    # one scene, passes alternating between two streams with their own
    # output buffers and no synchronisation in between: the scene keeps a
    # workspace per stream, so the overlapping passes must not disturb each
    # other; both buffer sets end with the host-output result
    gpu, _, m = icosphere
    n = 160
    cam = scenes.small_body_camera(m, resolution=n, seed=2)
    sun = scenes.sun_direction(3)
    ref = gpu.sun_pass(cam, sun)
    streams = [capi.Stream(), capi.Stream()]
    bufs = [(capi.DeviceBuffer(n * n * 8), capi.DeviceBuffer(n * n * 4), capi.DeviceBuffer(n * n)) for _ in streams]
    for b in bufs:
        for x in b:
            x.fill(0x33)
    outs = [{"depth": d.ptr, "tri": t.ptr, "lit": l.ptr} for d, t, l in bufs]
    for k in range(8):
        assert gpu.sun_pass(cam, sun, outs[k & 1], device=True, stream=streams[k & 1].handle, timing=False) is None
    capi.device_synchronize()
    for d, t, l in bufs:
        assert np.array_equal(d.to_host(np.empty((n, n))), ref["depth"])
        assert np.array_equal(t.to_host(np.empty((n, n), np.int32)), ref["tri"])
        assert np.array_equal(l.to_host(np.empty((n, n), np.uint8)), ref["lit"])


def test_empty_scene_rejected():
    with pytest.raises(ValueError):
        capi.Scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))


def test_bad_index_rejected():
    with pytest.raises(ValueError):
        capi.Scene(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32))


def test_single_triangle_scene():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    i = np.array([[0, 1, 2]], np.uint32)
    s = capi.Scene(v, i)
    t, tri, u, vv = s.trace(np.array([[0.25, 0.25, 1.0]]), np.array([[0.0, 0.0, -1.0]]))
    assert tri[0] == 0 and abs(t[0] - 1.0) < 1e-6
    assert abs(u[0] - 0.25) < 1e-6 and abs(vv[0] - 0.25) < 1e-6


def test_ploc_builder_matches(icosphere, cornell):
    # the PLOC binary tree (CRT_BUILD_PLOC) changes only the BVH: hits agree
    # with the LBVH scene (CRT_BUILD_LBVH) except at exact ties / edge grazes
    gpu, cpu, m = icosphere
    lbvh = capi.Scene(m.vertices, m.indices, ploc=False)
    alt = capi.Scene(m.vertices, m.indices, ploc=True)
    g_greedy = capi.Scene(m.vertices, m.indices, ploc=False, greedy_wide=True)
    a_greedy = capi.Scene(m.vertices, m.indices, ploc=True, greedy_wide=True)
    assert a_greedy.info()["sah_cost"] < g_greedy.info()["sah_cost"]
    cam = scenes.small_body_camera(m, resolution=512, seed=1)
    a, b = lbvh.camera_pass(cam), alt.camera_pass(cam)
    mism = a["tri"] != b["tri"]
    if mism.any():
        o, d = oracle.camera_rays(cam)
        sel = mism.ravel()
        ok = cpu.tie_flags(o[sel], d[sel]) | cpu.edge_flags(o[sel], d[sel], b["tri"].ravel()[sel])
        assert ok.all()
    hit = (~mism) & (a["tri"] >= 0)
    assert np.max(np.abs(a["depth"][hit] - b["depth"][hit]) / a["depth"][hit]) < REL
    # multi-instance scene
    g, c, v, i = cornell
    sc = scenes.cornell_box(0)
    _, _, inst = sc.merged()
    alt = capi.Scene(v, i, inst, ploc=False)   # the Cornell box's default tree is PLOC (36 triangles)
    cam = scenes.cornell_camera()
    pa, pc = alt.camera_pass(cam), c.camera_pass(cam)
    mism = pa["tri"] != pc["tri"]
    if mism.any():
        o, d = oracle.camera_rays(cam)
        sel = mism.ravel()
        ok = c.tie_flags(o[sel], d[sel]) | c.edge_flags(o[sel], d[sel], pa["tri"].ravel()[sel])
        assert ok.all()


def test_default_tree_by_size():
    # without a tree flag the library builds PLOC up to 131072 triangles and
    # LBVH above: the default scene's SAH is that of the explicit choice (the
    # SAH is a parallel sum, equal to ~1e-15) and not that of the other tree
    for level, ploc in ((6, True), (7, False)):
        m = scenes.displaced_icosphere(level)
        d = capi.Scene(m.vertices, m.indices).info()["sah_cost"]
        same = capi.Scene(m.vertices, m.indices, ploc=ploc).info()["sah_cost"]
        other = capi.Scene(m.vertices, m.indices, ploc=not ploc).info()["sah_cost"]
        assert abs(d - same) <= 1e-9 * same, level
        assert abs(d - other) > 1e-4 * other, level
    # the two tree flags exclude each other
    with pytest.raises(ValueError):
        opts = capi.BuildOpts(8, capi.BUILD_PLOC | capi.BUILD_LBVH)
        h = capi.C.c_void_p()
        v = np.ascontiguousarray(m.vertices, dtype=np.float32)
        i = np.ascontiguousarray(m.indices, dtype=np.uint32)
        capi._check(capi.load().crt_scene_create(capi.ptr(v), v.shape[0], capi.ptr(i), i.shape[0], None, None, 1,
                                                 None, None, capi.C.byref(opts), capi.C.byref(h)))


@pytest.mark.parametrize("ploc", [False, True])
def test_greedy_collapse_matches(icosphere, cornell, ploc):
    # the greedy collapse (CRT_BUILD_GREEDY_WIDE) and the default SAH-optimal
    # (DP) collapse change only the wide tree: closest hits, sun shadows and
    # any-hit queries agree except at exact ties / edge grazes, and every
    # triangle stays reachable
    gpu, cpu, m = icosphere
    alt = capi.Scene(m.vertices, m.indices, ploc=ploc, greedy_wide=True)
    inf = alt.info()
    assert inf["n_wide_nodes"] >= 1 and inf["max_depth"] >= 1
    cam = scenes.small_body_camera(m, resolution=512, seed=1)
    a, b = gpu.camera_pass(cam), alt.camera_pass(cam)
    mism = a["tri"] != b["tri"]
    if mism.any():
        o, d = oracle.camera_rays(cam)
        sel = mism.ravel()
        ok = cpu.tie_flags(o[sel], d[sel]) | cpu.edge_flags(o[sel], d[sel], b["tri"].ravel()[sel])
        assert ok.all()
    hit = (~mism) & (a["tri"] >= 0)
    assert np.max(np.abs(a["depth"][hit] - b["depth"][hit]) / a["depth"][hit]) < REL
    sun = scenes.sun_direction(2)
    sa, sb = gpu.sun_pass(cam, sun), alt.sun_pass(cam, sun)
    same = (sa["tri"] == sb["tri"]) & (sa["tri"] >= 0)
    assert np.mean(sa["lit"][same] != sb["lit"][same]) < 1e-3
    # any-hit queries from inside the body outwards: every ray is blocked
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(4096, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    t, tri, _, _ = alt.trace(np.zeros((4096, 3)), dirs, any_hit=True)
    assert np.all(tri >= 0)
    # multi-instance scene
    g, c, v, i = cornell
    sc = scenes.cornell_box(0)
    _, _, inst = sc.merged()
    alt = capi.Scene(v, i, inst, ploc=ploc, greedy_wide=True)
    cam = scenes.cornell_camera()
    pa, pc = alt.camera_pass(cam), c.camera_pass(cam)
    mism = pa["tri"] != pc["tri"]
    if mism.any():
        o, d = oracle.camera_rays(cam)
        sel = mism.ravel()
        ok = c.tie_flags(o[sel], d[sel]) | c.edge_flags(o[sel], d[sel], pa["tri"].ravel()[sel])
        assert ok.all()


def test_collapse_small_scenes():
    # 1, 2, 5, 9 and 17 triangles: roots that are a binary leaf, a single
    # leaf range, or need more than one wide level
    rng = np.random.default_rng(11)
    for nt in (1, 2, 5, 9, 17):
        v = rng.uniform(-1, 1, size=(3 * nt, 3)).astype(np.float32)
        idx = np.arange(3 * nt, dtype=np.uint32).reshape(nt, 3)
        a = capi.Scene(v, idx, greedy_wide=True)
        b = capi.Scene(v, idx)
        o = rng.uniform(-3, 3, size=(2048, 3))
        d = rng.normal(size=(2048, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        ta, ia, _, _ = a.trace(o, d)
        tb, ib, _, _ = b.trace(o, d)
        assert np.array_equal(ia >= 0, ib >= 0), nt
        both = (ia >= 0) & (ib >= 0)
        assert np.allclose(ta[both], tb[both], rtol=1e-6), nt
