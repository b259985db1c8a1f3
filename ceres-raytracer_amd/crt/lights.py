"""Light models (crt/lights.py of the reference) plus a directional sun."""
from abc import ABC

import numpy as np
from numpy.typing import ArrayLike

import _crt
from crt.rigid_body import RigidBody


class Light(ABC):
    intensity: float


class PointLight(RigidBody, Light):
    """Point light: irradiance min(I / d^2, 1e4) (src/crt/lights/point_light.hpp:36-39)."""

    def __init__(self, intensity: float, **kwargs):
        super(PointLight, self).__init__(**kwargs)
        self.intensity = intensity
        self._cpp = _crt.PointLight(self.intensity)
        self.set_pose(self.position, np.eye(3))


class AreaLight(RigidBody, Light):
    """Rectangular area light of side lengths `size` in its local XY plane;
    a uniformly sampled point is placed with R and position, irradiance
    I / d^2 (src/crt/lights/area_light.hpp:54-72)."""

    def __init__(self, intensity: float, size: ArrayLike, **kwargs):
        super(AreaLight, self).__init__(**kwargs)
        self.intensity = intensity
        self.size = size
        self._cpp = _crt.AreaLight(self.intensity, list(size))
        self.set_pose(self.position, self.rotation)


class DirectionalLight(RigidBody, Light):
    """Light at infinity (e.g. the sun): constant irradiance `intensity`
    arriving from `direction` (unit vector from the scene towards the light).
    New in this implementation; rotated into body frames like the others."""

    def __init__(self, intensity: float, direction: ArrayLike, **kwargs):
        super(DirectionalLight, self).__init__(**kwargs)
        self.intensity = intensity
        self.direction = np.asarray(direction, dtype=np.float64).reshape(3)
        self._cpp = _crt.DirectionalLight(self.intensity, self.direction)
        self.set_pose(self.position, self.rotation)
