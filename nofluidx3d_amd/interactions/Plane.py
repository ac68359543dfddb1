"""Flat indenter (reference ``interactions/Plane.py:18-41``, kernel ``Interactions/Plane.cl:1-7``):
every node above the plane y = y_p gets k * (y - y_p) along -y.

``wallFunc(time)`` gives y_p; as for :class:`Sphere`, a ``wallFunc.law`` attribute lets the fused
node pass compute y_p on the device.
"""
from nofluidx3d_amd import _lib
from nofluidx3d_amd.interactions.Interaction import DeviceCall, Interaction


class Plane(Interaction):
    def __init__(self, wallFunc, forceConst):
        super().__init__()
        self.wallFunc = wallFunc
        self.forceConst = forceConst
        self.height = None  # y_p of the step being enqueued

    def build(self):
        sim = self.simulation
        self.scope = sim.numPoints
        self.knl = DeviceCall("nfx_plane_force", lambda: (sim.force.ptr, sim.points.ptr, sim.numPoints,
                                                          float(self.forceConst), float(self.height)))

    def setArgs(self):
        self.height = self.wallFunc(self.simulation.time)

    def beforeTimeStep(self, globalTime):
        self.height = self.wallFunc(globalTime)

    def contribute(self, plan) -> bool:
        law = getattr(self.wallFunc, "law", None)
        if law is None or plan.indenter != _lib.INDENTER_NONE:
            return False
        plan.indenter, plan.ind_k, plan.ind_radius = _lib.INDENTER_PLANE, float(self.forceConst), 0.0
        plan.ind_start, plan.ind_speed, plan.ind_move_steps = map(float, law)
        return True
