"""Calls on one scene from several host threads: each scene's entry points
share its grow-only device workspaces behind a per-scene lock, so
concurrent calls must give exactly the results of the same calls made one
after another (and calls on different scenes run side by side)."""
import threading

import numpy as np
import pytest

from crt import capi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bodies():
    m = scenes.displaced_icosphere(6)
    return m, capi.Scene(m.vertices, m.indices), capi.Scene(m.vertices, m.indices)


def _jobs(m, s):
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    cams = [scenes.small_body_camera(m, resolution=r, seed=20 + k) for k, r in enumerate((96, 128, 64, 160))]
    return [
        lambda: s.camera_pass(cams[0])["depth"],
        lambda: s.render(cams[1], [sun], 2, 2, 0.0, 2, seed=3)[0],
        lambda: s.camera_pass(cams[2])["tri"],
        lambda: s.render(cams[3], [sun], 1, 1, 0.0, 3, seed=4)[0],
        lambda: np.concatenate(s.radiation_pressure(scenes.sun_direction(4), 4.56e-6, 0.3, 0.6, 256)[:2]),
    ]


def _run_threads(fns, repeat=3):
    out = [None] * (len(fns) * repeat)
    errs = []

    def work(i):
        try:
            out[i] = fns[i % len(fns)]()
        except Exception as e:   # pragma: no cover - reported below
            errs.append(e)

    ths = [threading.Thread(target=work, args=(i,)) for i in range(len(out))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errs, errs
    return out


def test_concurrent_calls_on_one_scene_match_sequential(bodies):
    m, s, _ = bodies
    fns = _jobs(m, s)
    ref = [f() for f in fns]
    got = _run_threads(fns)
    for i, g in enumerate(got):
        assert np.array_equal(g, ref[i % len(fns)]), i


def test_concurrent_calls_on_two_scenes(bodies):
    m, s1, s2 = bodies
    f1, f2 = _jobs(m, s1), _jobs(m, s2)
    ref = [f() for f in f1]
    got = _run_threads(f1 + f2, repeat=2)
    for i, g in enumerate(got):
        assert np.array_equal(g, ref[i % len(f1)]), i


def test_renders_on_separate_streams_match_sequential(bodies):
    # renders of one scene from three threads, each on its own stream (its
    # own workspace, so they run concurrently): every image equals the same
    # render made alone
    m, s, _ = bodies
    sun = {"type": "directional", "intensity": 1.0, "direction": scenes.sun_direction(2)}
    cams = [scenes.small_body_camera(m, resolution=r, seed=40 + k) for k, r in enumerate((96, 160, 128))]
    ref = [s.render_batch([c], [sun], 4, 4, 0.0, 3, seeds=[7 + k])[0][0][0] for k, c in enumerate(cams)]
    streams = [capi.Stream() for _ in cams]
    got = [None] * len(cams)

    def work(k):
        got[k] = s.render_batch([cams[k]], [sun], 4, 4, 0.0, 3, seeds=[7 + k], stream=streams[k].handle)[0][0][0]

    for _ in range(2):
        ths = [threading.Thread(target=work, args=(k,)) for k in range(len(cams))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        for k in range(len(cams)):
            assert np.array_equal(got[k], ref[k]), k
