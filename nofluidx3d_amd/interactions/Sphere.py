"""Rigid spherical indenter (reference ``interactions/Sphere.py:18-43``, kernel ``Sphere.cl``).

``sphereFunc(time)`` gives the sphere centre's y coordinate.  When the function carries a
``law = (start, speed_per_step, move_steps)`` attribute (``Simulation.setInteractionSphere`` sets
it), the fused step computes the position on the device from the step index instead of
re-setting a kernel argument every step.
"""
import ctypes
from pathlib import Path

from nofluidx3d_amd import _lib
from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.KernelBuilder import KernelBuilder

kernels = Path(__file__).parents[1] / "kernels"


class Sphere(Interaction):
    def __init__(self, radius, sphereFunc, forceConst):
        self.radius = radius
        self.sphereFunc = sphereFunc
        self.forceConst = forceConst

        self.globalTime = 0

    def build(self):
        self.scope = self.simulation.numPoints
        KernelBuilder.define(INSERT_NUM_POINTS=self.simulation.numPoints, def_FORCE_CONST=self.forceConst)
        self.knl = KernelBuilder.build(kernels / "Interactions" / "Sphere.cl", "Interaction_Sphere")

    def setArgs(self):
        self.knl.set_args(
            self.simulation.force.buf,
            self.simulation.points.buf,
            ctypes.c_double(self.radius),
            ctypes.c_double(self.sphereFunc(self.globalTime)),
        )

    def beforeTimeStep(self, globalTime):
        self.globalTime = globalTime
        self.setArgs()

    def contribute(self, plan) -> bool:
        law = getattr(self.sphereFunc, "law", None)
        if law is None or plan.indenter != _lib.INDENTER_NONE:
            return False
        plan.indenter = _lib.INDENTER_SPHERE
        plan.ind_k = float(self.knl.values["def_FORCE_CONST"])
        plan.ind_radius = float(self.radius)
        plan.ind_start, plan.ind_speed, plan.ind_move_steps = (float(v) for v in law)
        return True
