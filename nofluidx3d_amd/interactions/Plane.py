"""Flat indenter pushing along -y (reference ``interactions/Plane.py:18-41``, kernel ``Plane.cl``).

As for :class:`Sphere`, a ``wallFunc.law = (start, speed_per_step, move_steps)`` attribute lets
the fused step compute the plane position on the device.
"""
import ctypes
from pathlib import Path

from nofluidx3d_amd import _lib
from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.KernelBuilder import KernelBuilder

kernels = Path(__file__).parents[1] / "kernels"


class Plane(Interaction):
    def __init__(self, wallFunc, forceConst):
        self.wallFunc = wallFunc
        self.forceConst = forceConst

        self.globalTime = 0

    def build(self):
        self.scope = self.simulation.numPoints
        KernelBuilder.define(INSERT_NUM_POINTS=self.simulation.numPoints, def_FORCE_CONST=self.forceConst)
        self.knl = KernelBuilder.build(kernels / "Interactions" / "Plane.cl", "Interaction_Plane")

    def setArgs(self):
        self.knl.set_args(
            self.simulation.force.buf,
            self.simulation.points.buf,
            ctypes.c_double(self.wallFunc(self.globalTime)),
        )

    def beforeTimeStep(self, globalTime):
        self.globalTime = globalTime
        self.setArgs()

    def contribute(self, plan) -> bool:
        law = getattr(self.wallFunc, "law", None)
        if law is None or plan.indenter != _lib.INDENTER_NONE:
            return False
        plan.indenter = _lib.INDENTER_PLANE
        plan.ind_k = float(self.knl.values["def_FORCE_CONST"])
        plan.ind_radius = 0.0
        plan.ind_start, plan.ind_speed, plan.ind_move_steps = (float(v) for v in law)
        return True
