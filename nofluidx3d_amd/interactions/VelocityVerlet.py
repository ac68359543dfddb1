"""Damped velocity-Verlet integrator (reference ``interactions/VelocityVerlet.py:22-50``, kernel
``VelocityVerlet.cl``).  Register it last: it consumes and clears the accumulated force."""
from pathlib import Path

import numpy as np

from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.KernelBuilder import KernelBuilder

kernels = Path(__file__).parents[1] / "kernels"


class VelocityVerlet(Interaction):
    def __init__(self, numPoints, fixTopBottom=False):
        self.fixTopBottom = fixTopBottom

    def createBuffers(self):
        import torch

        numPoints = self.simulation.numPoints
        dev = self.simulation.points.buf.device
        # velocity and old force start at zero (VelocityVerlet.py:29-36)
        self.velocityB = torch.zeros(3 * numPoints, dtype=torch.float32, device=dev)
        self.forceOldB = torch.zeros(3 * numPoints, dtype=torch.float32, device=dev)

    def build(self):
        self.scope = self.simulation.numPoints
        KernelBuilder.define(INSERT_NUM_POINTS=self.simulation.numPoints)
        if self.fixTopBottom:
            KernelBuilder.define("FIXED_TOP_BOTTOM")
        self.knl = KernelBuilder.build(kernels / "Interactions" / "VelocityVerlet.cl", "VelocityVerlet")

    def setArgs(self):
        self.knl.set_args(self.simulation.points.buf, self.velocityB, self.simulation.force.buf, self.forceOldB)

    def contribute(self, plan) -> bool:
        plan.vel = self.velocityB.data_ptr()
        plan.fold = self.forceOldB.data_ptr()
        plan.fix_top_bottom = 1 if "FIXED_TOP_BOTTOM" in self.knl.defines else 0
        return True

    def velocities(self) -> np.ndarray:
        return self.velocityB.cpu().numpy().astype(np.float64)
