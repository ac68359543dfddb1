"""Interaction plugins (reference ``nofluidx3d/interactions/__init__.py:11-23``) plus the
finite-strain viscoplastic material and the legacy stress laws."""

from nofluidx3d_amd.interactions.FiniteStrainViscoplastic import FiniteStrainViscoplastic
from nofluidx3d_amd.interactions.LegacyElastic import (LegacyElastic, LinearElastic, LinearElasticDeviatoric,
                                                       LinearViscoelastic, PoroNeoHookean, SecondOrderNeoHookean)
from nofluidx3d_amd.interactions.Plane import Plane
from nofluidx3d_amd.interactions.Sphere import Sphere
from nofluidx3d_amd.interactions.Substrate import Substrate
from nofluidx3d_amd.interactions.Tetra import Tetra
from nofluidx3d_amd.interactions.VelocityVerlet import VelocityVerlet

__all__ = [
    "FiniteStrainViscoplastic",
    "LegacyElastic",
    "LinearElastic",
    "LinearElasticDeviatoric",
    "LinearViscoelastic",
    "PoroNeoHookean",
    "SecondOrderNeoHookean",
    "Plane",
    "Sphere",
    "Substrate",
    "Tetra",
    "VelocityVerlet",
]
