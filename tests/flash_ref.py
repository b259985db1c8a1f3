"""Flash-LiDAR sub-ray directions restated in FP64 numpy (test helper).

Boresight d of beam (pixel) b from the oracle camera, tangent frame
e1 = d x a (a = x axis, or y when |d.x| > 0.9), e2 = d x e1, two Gaussian
angles by Box-Muller from the counter RNG u01(seed, b, k, 0/1) of sub-ray
k = 1..K, direction d + tan(s g1) e1 + tan(s g2) e2 normalised -- the
definition the device kernel follows (ray 0 of a beam is its boresight).
"""
import math

import numpy as np

import oracle


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
    dirs = np.empty((nb, sub_rays, 3))
    for b in range(nb):
        for k in range(1, sub_rays + 1):
            r1, r2 = oracle.u01(seed, b, k, 0), oracle.u01(seed, b, k, 1)
            rad = math.sqrt(-2.0 * math.log(1.0 - r1))
            a1 = math.tan(sigma * rad * math.cos(2.0 * math.pi * r2))
            a2 = math.tan(sigma * rad * math.sin(2.0 * math.pi * r2))
            v = d[b] + a1 * e1[b] + a2 * e2[b]
            dirs[b, k - 1] = v / np.linalg.norm(v)
    return np.repeat(o[:, None, :], sub_rays, axis=1), dirs
