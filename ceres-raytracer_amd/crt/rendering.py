"""Dynamic-scene API: the BVH is rebuilt on the GPU for every call
(crt/rendering.py of the reference; src/crt/rendering_dynamic/render.hpp:19-62)."""
from typing import List, Tuple, Union

import numpy as np

import _crt
from crt._pybind_convert import validate_entities, validate_lights
from crt.cameras import Camera
from crt.entity import Entity
from crt.lidars import Lidar
from crt.lights import Light


def render(camera: Camera, lights: Union[Light, List[Light], Tuple[Light, ...]],
           entities: Union[Entity, List[Entity], Tuple[Entity, ...]],
           min_samples: int = 1, max_samples: int = 1, noise_threshold: float = 1., num_bounces: int = 1) -> np.ndarray:
    return _crt.render(camera._cpp, validate_lights(lights), validate_entities(entities), min_samples, max_samples,
                       noise_threshold, num_bounces)


def simulate_lidar(lidar: Lidar, entities: Union[Entity, List[Entity], Tuple[Entity, ...]], num_rays: int = 1):
    return _crt.simulate_lidar(lidar._cpp, validate_entities(entities), num_rays)


def normal_pass(camera: Camera, entities, return_image: bool = False):
    normals = _crt.normal_pass(camera._cpp, validate_entities(entities))
    if return_image:
        return normals, 255 * np.abs(normals)
    return normals


def intersection_pass(camera: Camera, entities, return_image: bool = False):
    xyz = _crt.intersection_pass(camera._cpp, validate_entities(entities))
    if return_image:
        image = np.sqrt(xyz[:, :, 0] ** 2 + xyz[:, :, 1] ** 2 + xyz[:, :, 2] ** 2)
        image = image - np.min(image)
        image = 255 * image / np.max(image)
        return xyz, image
    return xyz


def instance_pass(camera: Camera, entities, return_image: bool = False):
    ids = _crt.instance_pass(camera._cpp, validate_entities(entities))
    if return_image:
        unique_ids = np.unique(ids)
        colors = np.random.randint(0, high=255, size=(3, unique_ids.size))
        image = np.zeros((ids.shape[0], ids.shape[1], 3))
        for k, uid in enumerate(unique_ids):
            image[ids == uid, :] = colors[:, k]
        return ids, image
    return ids
