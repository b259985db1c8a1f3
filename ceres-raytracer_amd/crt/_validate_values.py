"""Argument validation for poses and scales (crt/_validate_values.py of the reference).

The reference wraps its checks in ``assert((cond, msg))`` tuples, which are
always true, so nothing is ever rejected there.  These versions perform the
checks their messages describe and raise ValueError.
"""
import numpy as np

ROTATION_TOL = 1e-9


def validate_scale(scale):
    if isinstance(scale, bool) or not isinstance(scale, (int, float, np.integer, np.floating)):
        raise ValueError("Scale must be defined as a single positive float/int")
    if not scale > 0:
        raise ValueError("Scale must be defined as a single positive float/int")
    return scale


def validate_position(position):
    arr = np.asarray(position, dtype=np.float64).squeeze()
    if arr.size != 3:
        raise ValueError("Positions must be a list/tuple/ndarray of 3 floats/ints")
    if not np.all(np.isfinite(arr)):
        raise ValueError("Positions must be finite")
    return np.array(arr.reshape(3), dtype=np.float64)


def validate_rotation(rotation):
    arr = np.asarray(rotation, dtype=np.float64)
    if arr.shape != (3, 3):
        raise ValueError("Rotations must be defined as ndarray with shape (3,3)")
    if abs(np.linalg.det(arr) - 1.0) > ROTATION_TOL * 1e3 or \
            np.linalg.norm(arr.T @ arr - np.eye(3)) > ROTATION_TOL * 1e3:
        raise ValueError("Rotation matrices must have determinant 1 and inverse equal to their transpose")
    return np.array(arr, dtype=np.float64)


def validate_positions(positions):
    """validate_position of every row of an (N, 3) array at once."""
    arr = np.asarray(positions, dtype=np.float64)
    if arr.ndim != 2 or arr.shape[1] != 3:
        raise ValueError("Positions must be an (N, 3) array of floats/ints")
    if not np.all(np.isfinite(arr)):
        raise ValueError("Positions must be finite")
    return np.array(arr, dtype=np.float64)


def validate_rotations(rotations):
    """validate_rotation of every R[:, :, k] of a (3, 3, N) array at once
    (the same determinant and orthogonality tolerances)."""
    arr = np.asarray(rotations, dtype=np.float64)
    if arr.ndim != 3 or arr.shape[:2] != (3, 3):
        raise ValueError("Rotations must be defined as ndarray with shape (3,3,N)")
    r = np.moveaxis(arr, 2, 0)                       # (N, 3, 3)
    det = np.linalg.det(r) if r.shape[0] else np.zeros(0)
    gram = np.einsum("kji,kjl->kil", r, r) - np.eye(3)
    bad = (np.abs(det - 1.0) > ROTATION_TOL * 1e3) | (np.sqrt(np.einsum("kij,kij->k", gram, gram)) > ROTATION_TOL * 1e3)
    if np.any(bad) or not np.all(np.isfinite(arr)):
        raise ValueError("Rotation matrices must have determinant 1 and inverse equal to their transpose")
    return np.array(arr, dtype=np.float64)
