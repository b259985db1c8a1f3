"""Conversion of Python light / entity arguments to the _crt objects."""
from crt.lights import AreaLight, DirectionalLight, PointLight


def valid_light(light):
    return isinstance(light, (PointLight, AreaLight, DirectionalLight))


def validate_lights(lights):
    items = lights if isinstance(lights, (list, tuple)) else [lights]
    out = []
    for light in items:
        if not valid_light(light):
            raise TypeError("lights must be PointLight, AreaLight or DirectionalLight")
        out.append(light._cpp)
    return out


def validate_entities(entities):
    from crt.entity import Entity
    items = entities if isinstance(entities, (list, tuple)) else [entities]
    out = []
    for e in items:
        if not isinstance(e, Entity):
            raise TypeError("entities must be crt.Entity objects")
        out.append(e._cpp)
    return out
