"""Materials on the GPU (src/crt/materials/material.hpp:50-60, 90-184):
closed forms the reference's formulas fix (Blinn-Phong with one term,
texture lookups, mirror reflection), same-key parity with the oracle on a
scene using every material and map, the drop-in `crt` API, and scene blobs
that keep their materials."""
import numpy as np
import pytest

import material_scenes as ms
import oracle
from crt import capi, scenes
from render_parity import classify, fp32_eps

pytestmark = pytest.mark.gpu


def _render(sc, spp=1, nb=1, seed=1, cam=None):
    s = ms.build(sc, gpu=True)
    rgba, _, _, _ = s.render(cam or sc["camera"], sc["lights"], spp, spp, 0.0, nb, seed=seed)
    return rgba[..., :3]


def test_blinn_phong_single_term_closed_forms():
    for mat, expect in (({"type": "blinn_phong", "coeffs": (1.0, 0.0, 0.0)}, lambda c: np.minimum(c, 1.0)),
                        ({"type": "blinn_phong", "coeffs": (0.0, 0.0, 0.6), "alpha": 0.0},
                         lambda c: np.full(3, 0.6, np.float32))):
        sc = ms.lit_plane(mat)
        rgb = _render(sc)
        hit = rgb.sum(-1) > 0
        want = ms.TWO_PI_F * expect(np.asarray(sc["colors"][0], np.float32)).astype(np.float32) * np.float32(0.8)
        assert hit.mean() > 0.5
        assert np.allclose(rgb[hit], want, rtol=2e-6, atol=0), (rgb[hit][:3], want)
    sc = ms.lit_plane({"type": "blinn_phong", "coeffs": (0.0, 1.0, 0.0)})
    bp = _render(sc)
    sc["materials"] = [{"type": "lambertian", "color": sc["colors"][0]}]
    assert np.array_equal(bp, _render(sc))


def test_texture_lookups_closed_forms():
    sc = ms.mixed_cornell()
    sc["textures"] = [np.full((3, 5, 4), 200, np.uint8)] + sc["textures"][1:]
    sc["materials"] = list(sc["materials"])
    cam = dict(sc["camera"], resolution=[64, 64])
    sc["materials"][2] = {"type": "textured_lambertian", "texture": 0}
    a = _render(sc, 4, 3, 2, cam)
    sc["materials"][2] = {"type": "lambertian", "color": (np.float32(200) / np.float32(255),) * 3}
    b = _render(sc, 4, 3, 2, cam)
    assert np.array_equal(a, b)
    sc = ms.checker_plane()
    rgb = _render(sc)
    want = ms.checker_expected(sc["camera"])
    sel = want >= 0
    assert sel.mean() > 0.5 and np.array_equal(rgb[..., 0][sel] > 0.5, want[sel] == 1)


def test_mirror_reflection_closed_form():
    sc = ms.mirror_room()
    rgb = _render(sc, 1, 2)
    want = ms.mirror_expected(sc["camera"])
    sel = want != 0
    assert sel.mean() > 0.3
    assert np.array_equal(np.sign(rgb[..., 0] - rgb[..., 1])[sel], want[sel])
    assert np.all(_render(sc, 1, 1)[sel] == 0)   # a mirror adds no direct light


def test_mixed_materials_match_oracle_same_keys():
    sc = ms.mixed_cornell()
    cam = dict(sc["camera"], resolution=[128, 128])
    gpu = ms.build(sc, gpu=True)
    cpu = ms.build(sc, gpu=False)
    info = gpu.info()
    rgba, _, _, _ = gpu.render(cam, sc["lights"], 16, 16, 0.0, 4, seed=5)
    rgb, _, frag = cpu.render(cam, sc["lights"], 16, 16, 0.0, 4, seed=5, offset=info["self_offset"],
                              eps=fp32_eps(info))
    r = classify(rgba[..., :3], rgb, frag)
    print("mixed materials: agree %.5f, differ %d (closest %d, shadow %d, texel %d, unexplained %d), "
          "image mean rel %.2e (%.2e agreeing)" % (r["agree"], r["differ"], r["closest"], r["shadow"], r["texel"],
                                                   r["unexplained"], r["mean_rel"], r["mean_rel_agreeing"]))
    assert r["unexplained"] == 0
    assert r["agree"] > 0.99
    assert r["mean_rel_agreeing"] < 1e-4
    # materials change the image (the mirror wall and the boxes differ from Lambertian)
    plain = capi.Scene(sc["vertices"], sc["indices"], sc["tri_inst"], colors=sc["colors"])
    p, _, _, _ = plain.render(cam, sc["lights"], 16, 16, 0.0, 4, seed=5)
    assert np.mean(np.any(p != rgba, axis=-1)) > 0.2


def test_materials_through_crt_api_and_blobs(tmp_path):
    from crt import Entity, materials, rendering
    from crt.cameras import SimpleCamera
    from crt.lights import PointLight
    sc = ms.mirror_room()
    paths = []
    for k, name in enumerate(("floor", "red", "green")):
        idx = sc["indices"][2 * k:2 * k + 2]
        m = scenes.Mesh(sc["vertices"][4 * k:4 * k + 4], (idx - 4 * k).astype(np.uint32))
        p = str(tmp_path / (name + ".obj"))
        scenes.write_obj(p, m)
        paths.append(p)
    ents = [Entity(paths[0], material=materials.MirrorMaterial()),
            Entity(paths[1], color=list(sc["colors"][1])), Entity(paths[2], color=list(sc["colors"][2]))]
    c = sc["camera"]
    cam = SimpleCamera(c["focal_length"], c["resolution"], c["sensor_size"], c["z_positive"],
                       position=c["position"], rotation=c["rotation"])
    L = sc["lights"][0]
    img = rendering.render(cam, PointLight(L["intensity"], position=L["position"]), ents, 1, 1, 1.0, 2)
    want = ms.mirror_expected(c)
    sel = want != 0
    got = np.sign(img[..., 0].astype(int) - img[..., 1].astype(int))
    assert np.array_equal(got[sel], want[sel])
    # a scene blob keeps materials, UVs and textures
    mc = ms.mixed_cornell()
    s = ms.build(mc, gpu=True)
    r = capi.Scene.from_bytes(s.to_bytes())
    cam2 = dict(mc["camera"], resolution=[48, 48])
    a = s.render(cam2, mc["lights"], 2, 2, 0.0, 3, seed=9)[0]
    b = r.render(cam2, mc["lights"], 2, 2, 0.0, 3, seed=9)[0]
    assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        s.set_materials([{"type": "mirror"}] * 2)   # one material per instance
