"""Batch pose handling on the host (no GPU): vectorised validation of many
poses and the batch body-frame transform (crt/body_fixed.py:87-101 of the
reference, applied to N poses at once)."""
import numpy as np
import pytest

from crt._validate_values import validate_rotations, validate_positions


def _rotations(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
                  np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
                  np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)
    return np.moveaxis(R, 0, 2)   # (3, 3, n)


def test_validate_many_poses():
    n = 200_000
    R = _rotations(n, 1)
    assert validate_rotations(R).shape == (3, 3, n)
    bad = R.copy()
    bad[:, :, 123_456] *= 1.001            # not orthonormal
    with pytest.raises(ValueError):
        validate_rotations(bad)
    refl = R.copy()
    refl[0, :, 99_999] *= -1.0             # determinant -1
    with pytest.raises(ValueError):
        validate_rotations(refl)
    P = np.random.default_rng(2).normal(size=(n, 3))
    assert np.array_equal(validate_positions(P), P)
    P[777, 1] = np.inf
    with pytest.raises(ValueError):
        validate_positions(P)


def test_batch_transform_to_body():
    from crt import BodyFixedGroup
    G = _rotations(1, 3)[:, :, 0]
    g = BodyFixedGroup.__new__(BodyFixedGroup)   # the transform needs only the pose and scale
    g.position = np.array([1.5e6, -2.0e5, 7.0e4])
    g.rotation = G
    g.scale = 2.5
    n = 50_000
    P = np.random.default_rng(4).normal(size=(n, 3)) * 1e4 + g.position
    R = _rotations(n, 5)
    rp, rr = g.transform_to_body_batch(P, R)
    # p' = s G (p - p_g), R' = (G R^T)^T by matrix products
    ep = g.scale * (P - g.position) @ G.T
    er = np.einsum("ijk,lj->ilk", R, G)
    assert np.max(np.abs(rp - ep)) <= 1e-15 * g.scale * 4e4
    assert np.max(np.abs(rr - er)) <= 4e-16
    # closed form: a pose with the group's own rotation maps to the identity,
    # and the group's position maps to the origin
    p0, r0 = g.transform_to_body(g.position, G)
    assert np.array_equal(p0, np.zeros(3))
    assert np.max(np.abs(r0 - np.eye(3))) < 1e-15
    # pose k alone gets the bits it gets in the batch
    for k in (0, 31_337, n - 1):
        pk, rk = g.transform_to_body(P[k], R[:, :, k])
        assert np.array_equal(pk, rp[k]) and np.array_equal(rk, rr[:, :, k])
