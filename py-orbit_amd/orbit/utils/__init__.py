"""Small helpers shared by the lattice classes (reference: py/orbit/utils)."""
from orbit.utils.orbitFinalize import orbitFinalize, OrbitFinalizeError
from orbit.utils.objects import NamedObject, TypedObject, ParamsDictObject

speed_of_light = 2.99792458e+8

__all__ = ["orbitFinalize", "OrbitFinalizeError", "NamedObject", "TypedObject",
           "ParamsDictObject", "speed_of_light"]
