"""Entity: a mesh with colour and pose for the dynamic (per-call BVH) API."""
import numpy as np
from numpy.typing import ArrayLike

import _crt
from crt.rigid_body import RigidBody


class Entity(RigidBody):
    """Mesh loaded at construction from an OBJ file (crt/entity.py of the reference).

    :param geometry_path: path of the mesh file
    :param color: RGB albedo (default [1,1,1])
    :param geometry_type: "obj" (the only supported format, like the reference)
    :param smooth_shading: interpolate the OBJ vertex normals
    :param material: (additive) a crt.materials material, default ColoredLambertian
    :param uv: (additive) per-vertex texture coordinates (n_vertices, 2)
    """

    def __init__(self, geometry_path: str, color: ArrayLike = [1, 1, 1], geometry_type: str = "obj",
                 smooth_shading: bool = False, material=None, uv=None, **kwargs):
        super(Entity, self).__init__(**kwargs)
        self.geometry_path = geometry_path
        self.geometry_type = geometry_type
        self.color = color
        self.smooth_shading = smooth_shading
        self._cpp = _crt.Entity(self.geometry_path, self.geometry_type, self.smooth_shading, list(self.color))
        self.set_pose(self.position, self.rotation)
        self.set_scale(self.scale)
        # additive: a crt.materials material (the reference hard-codes
        # ColoredLambertian) and per-vertex UVs for its texture maps
        self.material = None
        if material is not None or uv is not None:
            from crt import materials
            materials.apply(self, material, uv)
