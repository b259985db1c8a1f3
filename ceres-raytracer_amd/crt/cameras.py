"""Camera models (crt/cameras.py of the reference)."""
from abc import ABC, abstractmethod

import numpy as np
from numpy.typing import ArrayLike

import _crt
from crt.rigid_body import RigidBody


class Camera(ABC):
    @abstractmethod
    def get_fov(self, degrees=True):
        return


class SimpleCamera(RigidBody, Camera):
    """Pinhole camera.  Rays go through integer pixel coordinates with the
    boresight along -z (or +z if z_positive), rotated by R^T
    (src/crt/cameras/simple_camera.hpp:61-81).

    :param focal_length: focal length
    :param resolution: [W, H] in pixels
    :param sensor_size: [width, height] of the sensor, same units as focal_length
    :param z_positive: boresight along +z instead of -z
    """

    def __init__(self, focal_length: float, resolution: ArrayLike, sensor_size: ArrayLike,
                 z_positive: bool = False, **kwargs):
        super(SimpleCamera, self).__init__(**kwargs)
        self.focal_length = focal_length
        self.resolution = resolution
        self.sensor_size = sensor_size
        self.z_positive = z_positive
        self._cpp = _crt.SimpleCamera(focal_length, list(resolution), list(sensor_size), z_positive)
        self.set_pose(self.position, self.rotation)

    def get_fov(self, degrees: bool = True) -> np.ndarray:
        """Field of view as the reference computes it: 2 atan(s / (4 f))
        (crt/cameras.py:67).  Ray generation itself spans 2 atan(s / (2 f))."""
        fov = 2 * np.arctan2(np.asarray(self.sensor_size) / 2, 2 * self.focal_length)
        return np.rad2deg(fov) if degrees else fov
