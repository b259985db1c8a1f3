"""Flash-LiDAR sub-ray directions restated in FP64 numpy (test helper).

Boresight d of beam (pixel) b from the oracle camera, tangent frame
e1 = d x a (a = x axis, or y when |d.x| > 0.9), e2 = d x e1, two Gaussian
angles by Box-Muller from the counter RNG u01(seed, b, k, 0/1) of sub-ray
k = 1..K, direction d + tan(s g1) e1 + tan(s g2) e2 normalised -- the
definition the device kernel follows (ray 0 of a beam is its boresight).
"""
import numpy as np

import oracle

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def u01(seed, a, b, c):
    """Vectorised counter RNG (wrapping uint64 arithmetic), the same function
    as oracle.u01 (crt_oracle_u01); a, b, c broadcast."""
    with np.errstate(over="ignore"):
        a = np.asarray(a, dtype=np.uint64)
        b = np.asarray(b, dtype=np.uint64)
        c = np.asarray(c, dtype=np.uint64)
        h = _mix64(np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15))
        h = _mix64(h ^ a)
        h = _mix64(h ^ (b + np.uint64(0x632BE59BD9B4E019)))
        h = _mix64(h ^ (c * np.uint64(0x8CB92BA72F3D8DD7) + np.uint64(1)))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def tangent_frame(d):
    ax = np.zeros_like(d)
    ax[:, 0] = 1.0
    big = np.abs(d[:, 0]) > 0.9
    ax[big] = [0.0, 1.0, 0.0]
    e1 = np.cross(d, ax)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(d, e1)
    return e1, e2


def flash_subrays(cam, seed, sub_rays, sigma):
    """(origins, dirs) of every beam's sub-rays 1..K, shape (beams, K, 3)."""
    o, d = oracle.camera_rays(cam)
    e1, e2 = tangent_frame(d)
    nb = d.shape[0]
    b = np.arange(nb, dtype=np.uint64)[:, None]
    k = np.arange(1, sub_rays + 1, dtype=np.uint64)[None, :]
    r1, r2 = u01(seed, b, k, 0), u01(seed, b, k, 1)
    rad = np.sqrt(-2.0 * np.log(1.0 - r1))
    a1 = np.tan(sigma * rad * np.cos(2.0 * np.pi * r2))[..., None]
    a2 = np.tan(sigma * rad * np.sin(2.0 * np.pi * r2))[..., None]
    v = d[:, None, :] + a1 * e1[:, None, :] + a2 * e2[:, None, :]
    dirs = v / np.linalg.norm(v, axis=2, keepdims=True)
    return np.repeat(o[:, None, :], sub_rays, axis=1), dirs
