"""Damped velocity-Verlet time integration (reference ``interactions/VelocityVerlet.py:22-50``,
kernel ``Interactions/VelocityVerlet.cl:1-36``): unit node mass, dt = 1, damping xi = 0.005;
``fixTopBottom`` pins x and z of nodes 3 and 4.  It consumes and clears the accumulated force, so
it is registered last.

State: the velocity and the previous step's force, 3N FP32 each, starting at zero.  The fused step
advances the same state in its node pass (``contribute``).
"""
import numpy as np

from nofluidx3d_amd.interactions.Interaction import DeviceCall, Interaction


class VelocityVerlet(Interaction):
    def __init__(self, numPoints, fixTopBottom=False):
        super().__init__()
        self.numPoints = numPoints
        self.fixTopBottom = bool(fixTopBottom)

    def createBuffers(self):
        import torch

        n3 = 3 * self.simulation.numPoints
        dev = self.simulation.device
        self.velocityB = torch.zeros(n3, dtype=torch.float32, device=dev)
        self.forceOldB = torch.zeros(n3, dtype=torch.float32, device=dev)

    def build(self):
        sim = self.simulation
        self.scope = sim.numPoints
        self.knl = DeviceCall("nfx_velocity_verlet", lambda: (sim.points.ptr, self.velocityB.data_ptr(), sim.force.ptr,
                                                              self.forceOldB.data_ptr(), sim.numPoints,
                                                              int(self.fixTopBottom)))

    def contribute(self, plan) -> bool:
        plan.vel, plan.fold = self.velocityB.data_ptr(), self.forceOldB.data_ptr()
        plan.fix_top_bottom = int(self.fixTopBottom)
        return True

    def state_buffers(self):
        return [self.velocityB, self.forceOldB]

    def velocities(self) -> np.ndarray:
        return self.velocityB.cpu().numpy().astype(np.float64)
