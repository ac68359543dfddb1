"""Legacy contacts and integrators behind the current plugin protocol (reference
``Interactions.py:35-65, 440-481, 601-691, 692-817, 818-860``; set up by
``SimulationLegacy.py:139-330, 550-617, 657-692, 884-940``).  Each plugin launches one libnfx kernel that
restates the reference's OpenCL kernel as written (see ``include/nfx.h``):

- ``RoundTip``: spherical tip with a cylindrical shaft, prefactor 0.1 (``RoundTip.cl``);
- ``SphereIntegral`` / ``TipIntegral``: exp(-k dis) n body force integrated per tetra with Bower's
  4-point rule (``InteractionSphereIntegral.cl``, ``InteractionTipIntegral.cl``);
- ``AdhesivePlane`` / ``AdhesivePlaneSurfaceIntegral``: Lennard-Jones wall on the surface points or
  integrated over the surface triangles (``InteractionAdhesivePlane*.cl``);
  ``HardAdhesivePlaneSurfaceIntegral``: the exp(-k d) - A exp(-d) wall integrated the same way
  (``InteractionHardAdhesivePlaneSurfaceIntegral.cl``);
- ``CalcCOM`` + ``COMForce`` and ``PointZeroForce``: springs that hold the cell in place;
- ``VelocityVerlet2`` (mass-weighted, ``velocityVerlet2.cl``) and ``OverDamped`` (``OverDamped.cl``):
  integrators, registered last.
"""

import numpy as np

from nofluidx3d_amd import _lib, openCL
from nofluidx3d_amd.interactions.Interaction import Interaction


class _PositionedTip(Interaction):
    def __init__(self, posFunc):
        self.posFunc = posFunc
        self.position = posFunc(0)

    def beforeTimeStep(self, globalTime):
        self.position = self.posFunc(globalTime)


class RoundTip(_PositionedTip):
    def __init__(self, radius, sphereFunc, forceConst):
        super().__init__(sphereFunc)
        self.radius, self.forceConst = radius, forceConst  # the kernel uses 0.1 (RoundTip.cl:15)

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_round_tip_force", openCL.handle(), s.force.ptr, s.points.ptr, s.numPoints, float(self.radius),
                  float(self.position))


class _VolumeIntegral(_PositionedTip):
    tip_shape = 0
    half_angle = 0.0
    tip_x = tip_z = 0.0

    def enqueue(self):
        s = self.simulation
        h = openCL.handle()
        args = (h, self.tip_shape, s.points.ptr, s.tetras.ptr, s.numPoints, s.numTetra, float(self.radius),
                float(self.half_angle), float(self.tip_x), float(self.position), float(self.tip_z),
                float(self.forceConst))
        if getattr(s, "scatter", "atomic") in ("csr", "tiled"):
            off, slots, corners = s.node_csr()
            _lib.call("nfx_volume_integral_force", *args, None, corners.data_ptr())
            _lib.call("nfx_gather_corners", h, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), s.numPoints,
                      s.force.ptr)
        else:
            _lib.call("nfx_volume_integral_force", *args, s.force.ptr, None)


class SphereIntegral(_VolumeIntegral):
    def __init__(self, radius, sphereFunc, forceConst):
        super().__init__(sphereFunc)
        self.radius, self.forceConst = radius, forceConst


class TipIntegral(_VolumeIntegral):
    tip_shape = 1

    def __init__(self, radius, halfAngle, sphereFunc, forceConst, spherePosX=0.0, spherePosZ=0.0):
        super().__init__(sphereFunc)
        self.radius, self.half_angle, self.forceConst = radius, halfAngle, forceConst
        self.tip_x, self.tip_z = spherePosX, spherePosZ


class AdhesivePlane(Interaction):
    def __init__(self, wallFunc, adhesionConst, distance, pointOnSurface):
        self.wallFunc, self.adhesionConst, self.distance = wallFunc, adhesionConst, distance
        self.pointOnSurface = np.ascontiguousarray(pointOnSurface, dtype=np.int32)
        self.wall = wallFunc(0)

    def createBuffers(self):
        import torch

        self.onSurfaceB = torch.from_numpy(self.pointOnSurface).to(self.simulation.points.buf.device)
        torch.cuda.synchronize()

    def beforeTimeStep(self, globalTime):
        self.wall = self.wallFunc(globalTime)

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_adhesive_plane_force", openCL.handle(), s.force.ptr, s.points.ptr, s.numPoints, float(self.wall),
                  float(self.adhesionConst), float(self.distance), self.onSurfaceB.data_ptr())


class AdhesivePlaneSurfaceIntegral(AdhesivePlane):
    def __init__(self, wallFunc, adhesionConst, distance, triangles):
        super().__init__(wallFunc, adhesionConst, distance, np.zeros(0, np.int32))
        self.triangles = np.asarray(triangles, dtype=np.uint32).reshape(-1, 3)

    def createBuffers(self):
        import torch

        soa = np.ascontiguousarray(self.triangles.reshape(-1, order="F").view(np.int32))
        self.trianglesB = torch.from_numpy(soa).to(self.simulation.points.buf.device)
        torch.cuda.synchronize()

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_adhesive_plane_surface_force", openCL.handle(), s.force.ptr, s.points.ptr,
                  self.trianglesB.data_ptr(), s.numPoints, len(self.triangles), float(self.wall),
                  float(self.adhesionConst), float(self.distance))


class HardAdhesivePlaneSurfaceIntegral(AdhesivePlaneSurfaceIntegral):
    """Reference ``InteractionHardAdhesivePlaneSurfaceIntegral`` (``Interactions.py:778-817``, kernel
    ``InteractionHardAdhesivePlaneSurfaceIntegral.cl``): the area density exp(-k d) - A exp(-d),
    d = y - wall (k = ``forceConst``, A = ``adhesionConst``), integrated over the surface triangles."""

    def __init__(self, wallFunc, adhesionConst, forceConst, triangles):
        super().__init__(wallFunc, adhesionConst, 0.0, triangles)
        self.forceConst = forceConst

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_hard_adhesive_plane_surface_force", openCL.handle(), s.force.ptr, s.points.ptr,
                  self.trianglesB.data_ptr(), s.numPoints, len(self.triangles), float(self.wall),
                  float(self.adhesionConst), float(self.forceConst))


class CalcCOM(Interaction):
    """com[0] = mean x, com[1] = mean z of the node positions (FP64 device tensor ``comB``)."""

    def createBuffers(self):
        import torch

        self.comB = torch.zeros(2, dtype=torch.float64, device=self.simulation.points.buf.device)

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_center_of_mass", openCL.handle(), s.points.ptr, s.numPoints, self.comB.data_ptr())


class COMForce(Interaction):
    def __init__(self, springConst, calcCOM):
        self.springConst, self.calcCOM = springConst, calcCOM

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_com_force", openCL.handle(), s.force.ptr, s.numPoints, float(self.springConst),
                  self.calcCOM.comB.data_ptr())


class PointZeroForce(Interaction):
    def __init__(self, springConst):
        self.springConst = springConst

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_point_zero_force", openCL.handle(), s.force.ptr, s.points.ptr, s.numPoints,
                  float(self.springConst))


class VelocityVerlet2(Interaction):
    """Mass-weighted damped velocity Verlet; ``masses``: per-node lumped masses (N)."""

    def __init__(self, masses, damping=0.1):
        self.massesNP = np.asarray(masses, dtype=np.float64)
        self.damping = damping

    def createBuffers(self):
        from nofluidx3d_amd.util.Bridge import Bridge

        n = self.simulation.numPoints
        self.masses = Bridge(self.massesNP, datatype=np.float64)
        self.velocity = Bridge(np.zeros(3 * n), datatype=np.float64)
        self.forceOld = Bridge(np.zeros(3 * n), datatype=np.float64)

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_velocity_verlet2", openCL.handle(), s.points.ptr, self.velocity.ptr, s.force.ptr,
                  self.forceOld.ptr, self.masses.ptr, s.numPoints, float(self.damping))


class OverDamped(Interaction):
    def __init__(self, point3fix=True):
        self.point3fix = bool(point3fix)

    def enqueue(self):
        s = self.simulation
        _lib.call("nfx_over_damped", openCL.handle(), s.points.ptr, s.force.ptr, s.numPoints, int(self.point3fix))


def lumped_masses(tetras, volumes, n_points, rho=1.0):
    """Per-node masses: a quarter of each tetra's mass to each of its corners (the intent of
    SimulationLegacy.py:912-919, whose array is sized by the tetra count; here by the node count)."""
    m = np.zeros(n_points)
    np.add.at(m, np.asarray(tetras).ravel(), np.repeat(0.25 * rho * np.abs(np.asarray(volumes)), 4))
    return m


def surface_triangles(tetras):
    """Boundary faces of a tetra mesh (faces that belong to exactly one tetra), oriented as in the
    tetra, (n, 3) node ids."""
    t = np.asarray(tetras)
    faces = np.concatenate([t[:, [1, 2, 3]], t[:, [0, 3, 2]], t[:, [0, 1, 3]], t[:, [0, 2, 1]]])
    key = np.sort(faces, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    return faces[cnt[inv.ravel()] == 1]
