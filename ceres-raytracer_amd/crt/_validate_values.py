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
