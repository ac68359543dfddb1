"""Rigid substrate under the cell (reference ``interactions/Substrate.py:18-37``, kernel
``Interactions/Substrate.cl:1-7``): a penalty force k * max(wall - y, 0) along +y on every node
below the wall plane y = wallPosition.

Per-plugin step: ``nfx_substrate_force`` on the absolute positions.  Fused step: the same term is
evaluated inside the node pass (``contribute``).
"""
from nofluidx3d_amd.interactions.Interaction import DeviceCall, Interaction


class Substrate(Interaction):
    def __init__(self, wallPosition, forceConst):
        super().__init__()
        self.wallPosition = wallPosition
        self.forceConst = forceConst

    def build(self):
        sim = self.simulation
        self.scope = sim.numPoints
        self.knl = DeviceCall("nfx_substrate_force", lambda: (sim.force.ptr, sim.points.ptr, sim.numPoints,
                                                              float(self.forceConst), float(self.wallPosition)))

    def contribute(self, plan) -> bool:
        if plan.substrate:  # one substrate per fused step
            return False
        plan.substrate, plan.substrate_k, plan.wall = 1, float(self.forceConst), float(self.wallPosition)
        return True
