"""Batched rendering (many cameras, one set of launches): each camera in the
batch is bit-for-bit its plain single-camera render with the same seed, and
the hipGraph / persistent-kernel variants change nothing but speed."""
import numpy as np
import pytest

from crt import capi, scenes

pytestmark = pytest.mark.gpu


def _cams(m, n, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        res = int(2 ** rng.uniform(4, 7))
        c = scenes.small_body_camera(m, resolution=res, seed=1000 + k, range_km=(20.0, 500.0))
        out.append(c)
    return out


@pytest.fixture(scope="module")
def body():
    m = scenes.displaced_icosphere(6)
    return m, capi.Scene(m.vertices, m.indices)


def test_batch_equals_plain_renders(body):
    m, s = body
    cams = _cams(m, 6)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    seeds = [11 + k for k in range(len(cams))]
    out, tm = s.render_batch(cams, [sun], 4, 4, 0.0, 3, seeds=seeds)
    for k, c in enumerate(cams):
        rgba, _, spp, _ = s.render(c, [sun], 4, 4, 0.0, 3, seed=seeds[k])
        assert np.array_equal(out[k][0], rgba), k
        assert np.all(out[k][1] == 4)
    assert tm["rays_traced"] > 0


@pytest.mark.parametrize("variant", [dict(use_graph=True), dict(persistent=True, grab=4, blocks_per_cu=4),
                                     dict(use_graph=True, persistent=True, grab=2, samples_per_pass=1),
                                     dict(samples_per_pass=1), dict(samples_per_pass=1, use_graph=True),
                                     dict(fixed_share=1), dict(fixed_share=1, rays_per_octet=3, rays_per_quad=3)])
def test_batch_variants_identical(body, variant):
    # samples_per_pass=1: every path folds its own sample where it ends (no
    # k_fold); the same bits as the k_fold of passes of several samples
    m, s = body
    cams = _cams(m, 5, seed=9)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    base, _ = s.render_batch(cams, [sun], 3, 3, 0.0, 3, seeds=list(range(5)))
    alt, _ = s.render_batch(cams, [sun], 3, 3, 0.0, 3, seeds=list(range(5)), **variant)
    for a, b in zip(base, alt):
        assert np.array_equal(a[0], b[0])


@pytest.mark.parametrize("variant", [dict(extend_mode=2), dict(extend_mode=3, samples_per_pass=2, use_graph=True),
                                     dict()])
def test_batch_traversal_forms_agree(body, variant):
    # lane and octet traversal visit nodes in different orders, so only rays
    # whose closest hit is an exact tie between two triangles may differ
    m, s = body
    cams = _cams(m, 5, seed=9)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    base, _ = s.render_batch(cams, [sun], 3, 3, 0.0, 3, seeds=list(range(5)), extend_mode=1)
    alt, _ = s.render_batch(cams, [sun], 3, 3, 0.0, 3, seeds=list(range(5)), **variant)
    diff = sum(int(np.any(a[0] != b[0], axis=-1).sum()) for a, b in zip(base, alt))
    total = sum(a[0].size // 4 for a in base)
    print("traversal forms: %d of %d pixels differ" % (diff, total))
    assert diff <= max(2, total // 1000)


def test_batch_per_camera_spp_and_adaptive(body):
    m, s = body
    cams = _cams(m, 4, seed=3)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    spp = [1, 2, 5, 8]
    out, _ = s.render_batch(cams, [sun], 1, 8, 0.0, 2, seeds=[1, 2, 3, 4], spp=spp)
    for k in range(4):
        assert np.all(out[k][1] == spp[k])
        single, _, sp1, _ = s.render(cams[k], [sun], spp[k], spp[k], 0.0, 2, seed=k + 1)
        assert np.array_equal(out[k][0], single)
    # adaptive: same stop decisions alone and batched
    out2, _ = s.render_batch(cams, [sun], 2, 6, 0.02, 2, seeds=[1, 2, 3, 4])
    for k in range(4):
        single, _, sp1, _ = s.render(cams[k], [sun], 2, 6, 0.02, 2, seed=k + 1)
        assert np.array_equal(out2[k][1], sp1)
        assert np.array_equal(out2[k][0], single)
    # and with one sample per pass, where each path folds its own sample and
    # sets its pixel's stop flag in the kernel where it ends
    out3, _ = s.render_batch(cams, [sun], 2, 6, 0.02, 2, seeds=[1, 2, 3, 4], samples_per_pass=1)
    stopped = 0
    for k in range(4):
        assert np.array_equal(out3[k][1], out2[k][1])
        assert np.array_equal(out3[k][0], out2[k][0])
        stopped += int((out3[k][1] < 6).sum())
    assert stopped > 0   # the threshold stops some pixels early


@pytest.mark.parametrize("extend_mode,fused", [(1, 0), (2, 0), (3, 0), (1, 2), (3, 2), (0, 0)])
@pytest.mark.parametrize("n_lights,bounces", [(1, 3), (2, 2), (0, 2), (1, 1)])
def test_fused_equals_separate_kernels(body, extend_mode, fused, n_lights, bounces):
    # the fused bounce kernels (trace + shade, shadow + accumulate, camera
    # rays made in the lane trace kernel) run the stage kernels' arithmetic
    # in the same order: images must be bit-identical
    m, s = body
    cams = _cams(m, 5, seed=21)
    lights = [{"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)},
              {"type": "point", "intensity": 4.0e11, "position": [3.0e5, -1.0e5, 2.0e5]}][:n_lights]
    kw = dict(seeds=list(range(5)), extend_mode=extend_mode)
    sep, tms = s.render_batch(cams, lights, 2, 2, 0.0, bounces, separate_kernels=1, **kw)
    fus, tmf = s.render_batch(cams, lights, 2, 2, 0.0, bounces, separate_kernels=fused, **kw)
    for a, b in zip(sep, fus):
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1], b[1])
    assert tms["rays_traced"] == tmf["rays_traced"]
    assert tms["shadow_rays"] == tmf["shadow_rays"]


def test_coherence_sort_same_images(body):
    # ordering the bounce queues by ray-origin cell changes only the order in
    # which paths are traced: every image is bit-identical, on the octet
    # (extend_mode 3) and the all-lane (extend_mode 2) bounce kernels
    m, s = body
    cams = _cams(m, 5, seed=9)
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    seeds = [21 + k for k in range(len(cams))]
    for em in (3, 2):
        ref, _ = s.render_batch(cams, [sun], 4, 4, 0.0, 3, seeds=seeds, extend_mode=em)
        out, tm = s.render_batch(cams, [sun], 4, 4, 0.0, 3, seeds=seeds, extend_mode=em, coherence_sort=1)
        for k in range(len(cams)):
            assert np.array_equal(out[k][0], ref[k][0]), (em, k)
        assert tm["rays_traced"] > 0
