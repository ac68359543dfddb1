"""Rigid substrate penalty wall pushing along +y (reference ``interactions/Substrate.py:18-37``,
kernel ``Substrate.cl``)."""
import ctypes
from pathlib import Path

from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.KernelBuilder import KernelBuilder

kernels = Path(__file__).parents[1] / "kernels"


class Substrate(Interaction):
    def __init__(self, wallPosition, forceConst):
        self.wallPosition = wallPosition
        self.forceConst = forceConst

    def build(self):
        self.scope = self.simulation.numPoints
        KernelBuilder.define(INSERT_NUM_POINTS=self.simulation.numPoints, def_FORCE_CONST=self.forceConst)
        self.knl = KernelBuilder.build(kernels / "Interactions" / "Substrate.cl", "Interaction_Substrate")

    def setArgs(self):
        self.knl.set_args(
            self.simulation.force.buf,
            self.simulation.points.buf,
            ctypes.c_double(self.wallPosition),
        )

    def contribute(self, plan) -> bool:
        if plan.substrate:
            return False
        plan.substrate = 1
        plan.substrate_k = float(self.knl.values["def_FORCE_CONST"])
        plan.wall = float(self.wallPosition)
        return True
