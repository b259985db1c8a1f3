"""RigidBody: pose (position, rotation), scale and optional SPICE ephemerides.

Mirrors crt/rigid_body.py of the reference.  spiceypy is imported only by the
spice_* methods, so the rest of the package works without it.
"""
from typing import Union

import numpy as np
from numpy.typing import ArrayLike

from crt._validate_values import validate_position, validate_rotation, validate_scale


def _spice():
    try:
        import spiceypy
    except ImportError as e:  # pragma: no cover - depends on the environment
        raise ImportError("spiceypy is required for SPICE-driven poses") from e
    return spiceypy


class RigidBody:
    """Superclass of entities, cameras, lights and lidars.

    :param position: position of the body (3,)
    :param rotation: 3x3 rotation matrix (R^T = R^-1, det R = 1)
    :param scale: scale of the body
    :param name, frame, origin, ref, abcorr: NAIF names used by the spice_* methods
    """

    def __init__(self, position: ArrayLike = np.zeros(3), rotation: ArrayLike = np.eye(3),
                 scale: Union[float, int] = 1, name: Union[str, None] = None, frame: Union[str, None] = None,
                 origin: str = "SSB", ref: Union[str, None] = "J2000", abcorr: Union[str, None] = "NONE"):
        self.scale = validate_scale(scale)
        self.position = validate_position(position)
        self.rotation = validate_rotation(rotation)
        self.name = name
        self.frame = frame
        self.origin = origin
        self.ref = ref
        self.abcorr = abcorr

    def set_scale(self, scale: float):
        self.scale = validate_scale(scale)
        self._cpp.set_scale(self.scale)

    def set_position(self, position: ArrayLike):
        self.position = validate_position(position)
        self._cpp.set_position(self.position)

    def set_rotation(self, rotation: ArrayLike):
        self.rotation = validate_rotation(rotation)
        self._cpp.set_rotation(self.rotation)

    def set_pose(self, position: ArrayLike, rotation: ArrayLike):
        self.position = validate_position(position)
        self.rotation = validate_rotation(rotation)
        self._cpp.set_pose(self.position, self.rotation)

    def spice_position(self, et: float):
        position, _ = _spice().spkpos(self.name, et, self.ref, self.abcorr, self.origin)
        self.set_position(position)

    def spice_rotation(self, et: float):
        self.set_rotation(_spice().pxform(self.ref, self.frame, et))

    def spice_pose(self, et: float):
        sp = _spice()
        position, _ = sp.spkpos(self.name, et, self.ref, self.abcorr, self.origin)
        self.set_pose(position, sp.pxform(self.ref, self.frame, et))
