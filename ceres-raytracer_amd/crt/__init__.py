"""crt -- Python API of the MI355X ray-tracing core (drop-in for ceres-raytracer's `crt`).

Mirrors crt/__init__.py of ceres-raytracer v0.6a4: exports Entity, RigidBody,
BodyFixedGroup and BodyFixedEntity.  The classes are resolved lazily so that
the helper modules (`crt.scenes`, `crt.capi`) import without the compiled
`_crt` host module.
"""
__all__ = ['Entity', 'RigidBody', 'BodyFixedGroup', 'BodyFixedEntity']


def __getattr__(name):
    if name == "RigidBody":
        from crt.rigid_body import RigidBody
        return RigidBody
    if name == "Entity":
        from crt.entity import Entity
        return Entity
    if name in ("BodyFixedGroup", "BodyFixedEntity"):
        from crt import body_fixed
        return getattr(body_fixed, name)
    raise AttributeError(name)
