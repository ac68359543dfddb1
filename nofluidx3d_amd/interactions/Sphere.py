"""Rigid spherical indenter (reference ``interactions/Sphere.py:18-43``, kernel
``Interactions/Sphere.cl:1-17``): a node at distance d < radius from the centre (0, y_c, 0) is
pushed out along the radius with k * (radius - d).

``sphereFunc(time)`` gives y_c.  Per-plugin step: ``beforeTimeStep`` samples it and
``nfx_sphere_force`` applies the term.  Fused step: when ``sphereFunc`` carries
``law = (start, speed_per_step, move_steps)`` (``Simulation.setInteractionSphere`` attaches it),
the node pass computes y_c on the device from the step counter, and the plugin contributes the
law instead of a per-step argument.
"""
from nofluidx3d_amd import _lib
from nofluidx3d_amd.interactions.Interaction import DeviceCall, Interaction


class Sphere(Interaction):
    def __init__(self, radius, sphereFunc, forceConst):
        super().__init__()
        self.radius = radius
        self.sphereFunc = sphereFunc
        self.forceConst = forceConst
        self.centre = None  # y_c of the step being enqueued

    def build(self):
        sim = self.simulation
        self.scope = sim.numPoints
        self.knl = DeviceCall("nfx_sphere_force", lambda: (sim.force.ptr, sim.points.ptr, sim.numPoints,
                                                           float(self.forceConst), float(self.radius),
                                                           float(self.centre)))

    def setArgs(self):
        self.centre = self.sphereFunc(self.simulation.time)

    def beforeTimeStep(self, globalTime):
        self.centre = self.sphereFunc(globalTime)

    def contribute(self, plan) -> bool:
        law = getattr(self.sphereFunc, "law", None)
        if law is None or plan.indenter != _lib.INDENTER_NONE:
            return False
        plan.indenter, plan.ind_k, plan.ind_radius = _lib.INDENTER_SPHERE, float(self.forceConst), float(self.radius)
        plan.ind_start, plan.ind_speed, plan.ind_move_steps = map(float, law)
        return True
