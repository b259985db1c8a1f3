"""Elementary rotation matrices and Euler sequences (crt/rotations.py of the reference).

rotation_k(a) is the frame rotation about axis k by angle a; a sequence
"313" multiplies the elementary matrices left to right.
"""
import numpy as np


def rotation_1(rot):
    c, s = np.cos(rot), np.sin(rot)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def rotation_2(rot):
    c, s = np.cos(rot), np.sin(rot)
    return np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])


def rotation_3(rot):
    c, s = np.cos(rot), np.sin(rot)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


_ELEMENTARY = {"1": rotation_1, "2": rotation_2, "3": rotation_3}


def euler_to_rotmat(sequence, angles, degrees=False):
    if degrees:
        angles = np.deg2rad(angles)
    out = np.eye(3)
    for axis, angle in zip(sequence, angles):
        if axis not in _ELEMENTARY:
            raise Exception("Not a valid euler angle sequence!")
        out = out @ _ELEMENTARY[axis](angle)
    return out
