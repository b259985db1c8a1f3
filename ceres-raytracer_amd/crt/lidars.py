"""LiDAR models (crt/lidars.py of the reference)."""
from abc import ABC

import numpy as np
from numpy.typing import ArrayLike

import _crt
from crt.rigid_body import RigidBody, _spice
from crt._validate_values import validate_positions, validate_rotation, validate_rotations


class Lidar(ABC):
    def batch_set_pose(self, batch_positions: ArrayLike, batch_rotations: ArrayLike):
        """Poses for batch_simulate_lidar: positions (N,3), rotations (3,3,N).

        The reference caps N at 100000 (src/crt/lidars/lidar.hpp:11); the HIP
        core has no cap."""
        batch_positions = np.asarray(batch_positions, dtype=np.float64).reshape(-1, 3)
        batch_rotations = np.asarray(batch_rotations, dtype=np.float64)
        n = batch_positions.shape[0]
        if batch_rotations.ndim != 3 or batch_rotations.shape[:2] != (3, 3) or batch_rotations.shape[2] < n:
            raise ValueError("batch_rotations must have shape (3,3,N)")
        self.batch_positions = validate_positions(batch_positions)
        rot = validate_rotations(batch_rotations[:, :, :n])
        self.batch_rotations = rot
        self._cpp.batch_set_pose(np.ascontiguousarray(self.batch_positions), np.ascontiguousarray(rot), n)

    def batch_spice_pose(self, ets: np.ndarray):
        sp = _spice()
        positions, _ = sp.spkpos(self.name, ets, self.ref, self.abcorr, self.origin)
        rotations = np.zeros((3, 3, np.size(ets)))
        for k, et in enumerate(np.atleast_1d(ets)):
            rotations[:, :, k] = validate_rotation(sp.pxform(self.ref, self.frame, et))
        self.batch_set_pose(np.asarray(positions).reshape(-1, 3), rotations)


class SimpleLidar(RigidBody, Lidar):
    """Single-beam LiDAR along -z (or +z if z_positive) rotated by R^T
    (src/crt/lidars/simple_lidar.hpp:42-64)."""

    def __init__(self, z_positive: bool = False, **kwargs):
        super(SimpleLidar, self).__init__(**kwargs)
        self.z_positive = z_positive
        self._cpp = _crt.SimpleLidar(z_positive)
        self.set_pose(self.position, self.rotation)
