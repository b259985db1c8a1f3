"""Traversal statistics (crt_trace_stats, the reference traverser's
Statistics): the counted traversal returns the hits of crt_trace_rays, and
the counts follow what the tree fixes."""
import numpy as np
import pytest

import oracle
from crt import capi, scenes

pytestmark = pytest.mark.gpu


def test_single_triangle_counts_are_exact():
    # one triangle: the root is the only wide node, with one valid child (a
    # one-triangle leaf); a ray aimed inside the triangle visits 1 node, tests
    # 1 box and 1 triangle, and hits
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    i = np.array([[0, 1, 2]], dtype=np.uint32)
    sc = capi.Scene(v, i)
    assert sc.info()["n_wide_nodes"] == 1
    rng = np.random.default_rng(1)
    n = 4096
    o = rng.normal(size=(n, 3)) * 4.0
    o[:, 2] = np.abs(o[:, 2]) + 0.5
    a, b = rng.uniform(0.05, 0.45, size=n), rng.uniform(0.05, 0.45, size=n)
    tgt = np.stack([a, b, np.zeros(n)], axis=1)
    st = sc.trace_stats(o, tgt - o)
    assert np.all(st["nodes"] == 1) and np.all(st["boxes"] == 1) and np.all(st["tris"] == 1)
    assert np.all(st["tri"] == 0)
    t, tri, _, _ = sc.trace(o, tgt - o)
    assert np.array_equal(tri, st["tri"]) and np.array_equal(t, st["t"])
    # rays that miss the scene's bounding sphere are not traversed at all
    far = sc.trace_stats(o + 100.0, np.tile([1.0, 0.0, 0.0], (n, 1)))
    assert np.all(far["nodes"] == 0) and np.all(far["tri"] == -1)


@pytest.mark.parametrize("any_hit", [False, True])
def test_counted_traversal_same_hits(any_hit):
    mesh = scenes.displaced_icosphere(6, 2000.0, 0.2, 6, 0)
    sc = capi.Scene(mesh.vertices, mesh.indices)
    cam = scenes.small_body_camera(mesh, resolution=256, seed=3)
    o, d = scenes.camera_rays(cam)
    # the numpy pinhole rays are the oracle's camera rays
    oo, od = oracle.camera_rays(cam)
    assert np.allclose(o, oo, rtol=0, atol=0) and np.allclose(d, od, rtol=1e-15, atol=1e-15)
    st = sc.trace_stats(o, d, any_hit=any_hit)
    t, tri, _, _ = sc.trace(o, d, any_hit=any_hit)
    assert np.array_equal(tri >= 0, st["tri"] >= 0)
    if not any_hit:
        assert np.array_equal(tri, st["tri"]) and np.array_equal(t, st["t"])
    hit = st["tri"] >= 0
    assert hit.sum() > 1000
    # a hit needs at least one node, one box and one triangle; boxes are at
    # most 8 per node; any hit never tests more than closest hit on average
    assert np.all(st["nodes"][hit] >= 1) and np.all(st["tris"][hit] >= 1)
    assert np.all(st["boxes"] <= 8 * st["nodes"]) and np.all(st["boxes"] >= st["nodes"])
    info = sc.info()
    assert np.all(st["nodes"] <= info["n_wide_nodes"])
    print("L6 256^2 %s: nodes %.2f boxes %.2f tris %.2f per ray" % (
        "any" if any_hit else "closest", st["nodes"].mean(), st["boxes"].mean(), st["tris"].mean()))
