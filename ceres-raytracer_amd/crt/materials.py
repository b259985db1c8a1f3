"""Materials for entities (additive API).

The reference defines ColoredLambertianMaterial, TexturedLambertianMaterial,
TexturedBlinnPhongMaterial and MirrorMaterial (src/crt/materials/
material.hpp:90-184) but its Entity hard-codes the first
(src/crt/rendering_dynamic/entity.hpp:51-53).  Here an Entity or
BodyFixedEntity takes `material=` one of these classes (same names and
constructor arguments) and the GPU path tracer shades with it.

Texture maps are images (a PNG path, read like the reference's lodepng-based
ImageUVMap, or an (H, W, 3|4) uint8 array); lookups are the reference's
nearest texel at ((size_t)(u W + 0.5), (size_t)(v H + 0.5)) of the
interpolated per-vertex UV, clamped to the image.  The reference's OBJ
loader sets no UVs, so `uv=` (n_vertices x 2) on the entity supplies them.
"""
from __future__ import annotations

import struct
import zlib
from typing import Sequence, Union

import numpy as np

Image = Union[str, np.ndarray]


def load_png(path: str) -> np.ndarray:
    """Decode an 8-bit PNG (grey, grey+alpha, RGB, RGBA or palette; not
    interlaced) into an (H, W, 4) uint8 RGBA array."""
    with open(path, "rb") as fh:
        data = fh.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("%s: not a PNG file" % path)
    pos, idat, plte, trns = 8, [], None, None
    width = height = depth = ctype = interlace = None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            width, height, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            plte = np.frombuffer(body, dtype=np.uint8).reshape(-1, 3)
        elif kind == b"tRNS":
            trns = np.frombuffer(body, dtype=np.uint8)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if width is None or depth != 8 or interlace != 0 or ctype not in (0, 2, 3, 4, 6):
        raise ValueError("%s: only 8-bit, non-interlaced PNG images are supported" % path)
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), dtype=np.uint8).reshape(height, 1 + width * ch)
    out = np.zeros((height, width * ch), dtype=np.int32)
    prev = np.zeros(width * ch, dtype=np.int32)
    for y in range(height):
        f, line = raw[y, 0], raw[y, 1:].astype(np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 255
        else:
            cur = np.zeros_like(line)
            for x in range(line.size):
                a = cur[x - ch] if x >= ch else 0
                b = prev[x]
                c = prev[x - ch] if x >= ch else 0
                if f == 1:
                    p = a
                elif f == 3:
                    p = (a + b) >> 1
                else:   # Paeth
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[x] = (line[x] + p) & 255
        out[y] = cur
        prev = cur
    px = out.reshape(height, width, ch).astype(np.uint8)
    rgba = np.full((height, width, 4), 255, dtype=np.uint8)
    if ctype == 3:
        rgba[..., :3] = plte[px[..., 0]]
        if trns is not None:
            alpha = np.full(plte.shape[0], 255, dtype=np.uint8)
            alpha[:trns.size] = trns
            rgba[..., 3] = alpha[px[..., 0]]
    elif ctype in (0, 4):
        rgba[..., :3] = px[..., :1]
        if ctype == 4:
            rgba[..., 3] = px[..., 1]
    else:
        rgba[..., :ch] = px
    return rgba


def _image(img: Image) -> np.ndarray:
    if isinstance(img, str):
        return load_png(img)
    a = np.asarray(img)
    if a.dtype != np.uint8 or a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError("a texture must be a PNG path or an (H, W, 3|4) uint8 array")
    return a


class Material:
    kind = "lambertian"

    def _apply(self, cpp_entity):
        raise NotImplementedError


class ColoredLambertianMaterial(Material):
    """colour * (-L.n) (material.hpp:90-111); `color` None keeps the entity colour."""

    def __init__(self, color: Union[Sequence[float], None] = None):
        self.color = None if color is None else [float(c) for c in color]

    def _apply(self, cpp_entity):
        cpp_entity.set_material("lambertian")


class TexturedLambertianMaterial(Material):
    """texture(uv) * (-L.n) (material.hpp:113-136)."""

    def __init__(self, texture: Image):
        self.texture = _image(texture)

    def _apply(self, cpp_entity):
        cpp_entity.set_material("textured_lambertian", self.texture)


class TexturedBlinnPhongMaterial(Material):
    """clamp(texture(uv) (ka + kd diffuse) + ks pow(n.h, alpha)) with (ka, kd, ks)
    from the R, G, B of `specular` (an image, or three constants)
    (material.hpp:138-168)."""

    def __init__(self, texture: Image, specular: Union[Image, Sequence[float]], alpha: float = 24.0):
        self.texture = None if texture is None else _image(texture)
        if isinstance(specular, str) or (isinstance(specular, np.ndarray) and specular.ndim == 3):
            self.specular = _image(specular)
        else:
            self.specular = [float(x) for x in specular]
        self.alpha = float(alpha)

    def _apply(self, cpp_entity):
        cpp_entity.set_material("blinn_phong", self.texture, self.specular, self.alpha)


class MirrorMaterial(Material):
    """No direct light; the path continues along d - 2 n (d.n) (material.hpp:170-184)."""

    def _apply(self, cpp_entity):
        cpp_entity.set_material("mirror")


def apply(entity, material: Union[Material, None], uv=None):
    """Set `material` and per-vertex `uv` on an Entity / BodyFixedEntity."""
    if material is not None:
        if not isinstance(material, Material):
            raise TypeError("material must be a crt.materials.Material")
        if isinstance(material, ColoredLambertianMaterial) and material.color is not None:
            raise ValueError("ColoredLambertianMaterial takes the entity's colour (pass color= to the entity)")
        material._apply(entity._cpp)
        entity.material = material
    if uv is not None:
        entity._cpp.set_uv(np.ascontiguousarray(uv, dtype=np.float64))
