"""Body-fixed rendering: a group of entities whose BVH is built once on the GPU
and stays resident in HBM (crt/body_fixed.py of the reference)."""
from typing import List, Tuple, Union

import numpy as np
from numpy.typing import ArrayLike

import _crt
from crt.cameras import Camera
from crt.lidars import Lidar
from crt.lights import Light
from crt.rigid_body import RigidBody


def _to_body(group, positions, rotations):
    """p' = s G (p - p_g) and R' = (G R^T)^T = R G^T for N poses (positions
    (N, 3), rotations (3, 3, N)), every product summed left to right over j
    in FP64 so that one pose and a batch give the same bits."""
    G = np.asarray(group.rotation, dtype=np.float64)
    d = np.asarray(positions, dtype=np.float64) - np.asarray(group.position, dtype=np.float64)
    rp = np.empty_like(d)
    for i in range(3):
        rp[:, i] = group.scale * (G[i, 0] * d[:, 0] + G[i, 1] * d[:, 1] + G[i, 2] * d[:, 2])
    R = np.asarray(rotations, dtype=np.float64)
    rr = np.empty_like(R)
    for i in range(3):   # R'[i, l] = sum_j R[i, j] G[l, j]
        for l in range(3):
            rr[i, l, :] = R[i, 0, :] * G[l, 0] + R[i, 1, :] * G[l, 1] + R[i, 2, :] * G[l, 2]
    return rp, rr


class BodyFixedEntity(RigidBody):
    """Mesh, colour and pose of one member of a BodyFixedGroup.

    :param geometry_path: path of the OBJ mesh
    :param color: RGB albedo (default [1,1,1])
    :param geometry_type: "obj"
    :param smooth_shading: interpolate the OBJ vertex normals
    :param material: (additive) a crt.materials material, default ColoredLambertian
    :param uv: (additive) per-vertex texture coordinates (n_vertices, 2)
    """

    def __init__(self, geometry_path: str, color: ArrayLike = [1, 1, 1], geometry_type: str = "obj",
                 smooth_shading: bool = False, material=None, uv=None, **kwargs):
        super(BodyFixedEntity, self).__init__(**kwargs)
        self.geometry_path = geometry_path
        self.geometry_type = geometry_type
        self.color = color
        self.smooth_shading = smooth_shading
        self._cpp = _crt.BodyFixedEntity(self.geometry_path, self.geometry_type, self.smooth_shading,
                                         list(self.color))
        self.set_pose(self.position, self.rotation)
        self.set_scale(self.scale)
        # additive: a crt.materials material (the reference hard-codes
        # ColoredLambertian) and per-vertex UVs for its texture maps
        self.material = None
        if material is not None or uv is not None:
            from crt import materials
            materials.apply(self, material, uv)


class BodyFixedGroup(RigidBody):
    """Entities rendered in the body frame with a cached BVH.

    All methods first move the camera / lights / lidar into the body frame
    with transform_to_body and -- like the reference -- store that body-frame
    pose back into the object passed in (crt/body_fixed.py:128-140).

    :param entities: BodyFixedEntity or a list/tuple of them (instance ids
        1..N in order)
    """

    def __init__(self, entities: Union[BodyFixedEntity, List[BodyFixedEntity], Tuple[BodyFixedEntity, ...]],
                 **kwargs):
        super(BodyFixedGroup, self).__init__(**kwargs)
        items = entities if isinstance(entities, (list, tuple)) else [entities]
        for e in items:
            if not isinstance(e, BodyFixedEntity):
                raise TypeError("BodyFixedGroup expects BodyFixedEntity objects")
        self.entities = list(items)
        self._cpp = _crt.BodyFixedGroup([e._cpp for e in items])

    def transform_to_body(self, position: ArrayLike, rotation: ArrayLike) -> Tuple[np.ndarray, np.ndarray]:
        """FP64 pose in the group frame: p' = scale R_g (p - p_g), R' = (R_g R^T)^T
        (crt/body_fixed.py:87-101; the reference multiplies by the scale)."""
        p, r = _to_body(self, np.asarray(position, dtype=np.float64).reshape(1, 3),
                        np.asarray(rotation, dtype=np.float64).reshape(3, 3, 1))
        return p[0], r[:, :, 0]

    def transform_to_body_batch(self, positions: ArrayLike, rotations: ArrayLike) -> Tuple[np.ndarray, np.ndarray]:
        """transform_to_body of N poses at once: positions (N, 3), rotations
        (3, 3, N); pose k gets the bits transform_to_body gives it alone."""
        return _to_body(self, positions, rotations)

    def _to_body(self, obj):
        p, r = self.transform_to_body(obj.position, obj.rotation)
        obj.set_pose(p, r)

    def _lights(self, lights):
        items = lights if isinstance(lights, (list, tuple)) else [lights]
        out = []
        for light in items:
            self._to_body(light)
            if hasattr(light, "direction"):
                light._cpp = _crt.DirectionalLight(light.intensity, self.rotation @ light.direction)
            out.append(light._cpp)
        return out

    def render(self, camera: Camera, lights: Union[Light, List[Light], Tuple[Light, ...]],
               min_samples: int = 1, max_samples: int = 1, noise_threshold: float = 1., num_bounces: int = 1,
               seed: Union[int, None] = None) -> np.ndarray:
        """Path-traced image, uint8 (H, W, 4) RGBA like the reference.  `seed`
        (additive) fixes the counter-based RNG; None draws a fresh one."""
        self._to_body(camera)
        lights_cpp = self._lights(lights)
        return self._cpp.render(camera._cpp, lights_cpp, min_samples, max_samples, noise_threshold, num_bounces,
                                seed)

    def render_radiance(self, camera: Camera, lights, min_samples: int = 1, max_samples: int = 1,
                        noise_threshold: float = 1., num_bounces: int = 1, seed: Union[int, None] = None):
        """Float radiance (H, W, 4) f32 before the uint8 conversion (additive)."""
        self._to_body(camera)
        lights_cpp = self._lights(lights)
        return self._cpp.render_radiance(camera._cpp, lights_cpp, min_samples, max_samples, noise_threshold,
                                         num_bounces, seed)

    def simulate_lidar(self, lidar: Lidar, num_rays: int = 1):
        self._to_body(lidar)
        return self._cpp.simulate_lidar(lidar._cpp, num_rays)

    def batch_simulate_lidar(self, lidar: Lidar, num_rays: int = 1):
        rel_p, rel_r = self.transform_to_body_batch(lidar.batch_positions, lidar.batch_rotations)
        lidar.batch_set_pose(rel_p, rel_r)
        return self._cpp.batch_simulate_lidar(lidar._cpp, num_rays)

    def normal_pass(self, camera: Camera, return_image: bool = False):
        self._to_body(camera)
        normals = self._cpp.normal_pass(camera._cpp)
        if return_image:
            return normals, 255 * np.abs(normals)
        return normals

    def intersection_pass(self, camera: Camera, return_image: bool = False):
        self._to_body(camera)
        xyz = self._cpp.intersection_pass(camera._cpp)
        if return_image:
            image = np.sqrt(xyz[:, :, 0] ** 2 + xyz[:, :, 1] ** 2 + xyz[:, :, 2] ** 2)
            image = image - np.min(image)
            image = 255 * image / np.max(image)
            return xyz, image
        return xyz

    def instance_pass(self, camera: Camera, return_image: bool = False):
        self._to_body(camera)
        ids = self._cpp.instance_pass(camera._cpp)
        if return_image:
            unique_ids = np.unique(ids)
            colors = np.random.randint(0, high=255, size=(3, unique_ids.size))
            image = np.zeros((ids.shape[0], ids.shape[1], 3))
            for k, uid in enumerate(unique_ids):
                image[ids == uid, :] = colors[:, k]
            return ids, image
        return ids

    # ---- additive API -------------------------------------------------------
    def depth_pass(self, camera: Camera) -> np.ndarray:
        """Distance camera -> hit point (H, W) f64, 0 on a miss."""
        self._to_body(camera)
        return self._cpp.depth_pass(camera._cpp)

    def triangle_pass(self, camera: Camera) -> np.ndarray:
        """Index of the hit triangle in group order (H, W) int32, -1 on a miss."""
        self._to_body(camera)
        return self._cpp.triangle_pass(camera._cpp)

    def sun_pass(self, camera: Camera, sun_direction: ArrayLike):
        """Depth (H, W) f64 and sunlit mask (H, W) uint8 for a directional sun
        (unit vector towards the sun, base frame)."""
        self._to_body(camera)
        return self._cpp.sun_pass(camera._cpp, self.rotation @ np.asarray(sun_direction, dtype=np.float64))

    def flash_lidar(self, beams: Camera, sub_rays: int = 8, sigma: float = 5e-4, seed: int = 0) -> dict:
        """Flash LiDAR over the beam grid of `beams` (a SimpleCamera): single /
        first / last returns (H, W) f64 and the number of sub-ray returns."""
        self._to_body(beams)
        return self._cpp.flash_lidar(beams._cpp, sub_rays, sigma, seed)

    def radiation_pressure(self, sun_direction: ArrayLike, pressure: float, rho_s: float, rho_d: float,
                           grid: int = 2048, ref_point: ArrayLike = (0.0, 0.0, 0.0)):
        """Solar radiation force and torque (body frame, about `ref_point` in the
        body frame) from a grid x grid parallel ray lattice; `pressure` = flux / c."""
        sun_body = self.rotation @ np.asarray(sun_direction, dtype=np.float64)
        f, t, hits = self._cpp.radiation_pressure(sun_body, pressure, rho_s, rho_d, grid,
                                                  np.asarray(ref_point, dtype=np.float64))
        return f, t

    def bvh_info(self) -> dict:
        return self._cpp.bvh_info()
