"""AFM indentation simulation of one tetrahedral cell (reference ``nofluidx3d/Simulation.py``).

Same constructor, setters, ``register``, ``timeStep`` and ``run`` as the reference, the same JSON
config keys and unit system (``Simulation.py:42-56``), and the same recorded quantities
(time, indentation [m], tip force [N]) written to ``simdata.dat``.

Stepping.  When the registered plugins form the reference's standard chain -- one force plugin
(Tetra or FiniteStrainViscoplastic), optionally a Substrate and one indenter (Sphere or Plane)
whose position law is known, and VelocityVerlet -- ``timeStep`` and ``run`` advance with the
fused device step ``nfx_run_steps`` (a force pass and one node pass per step, indenter position
computed on the device, replayed from a captured hipGraph of up to 100 steps).  Any other chain
runs plugin by plugin, exactly like the reference's loop (``Simulation.py:232-237``).

Extensions (keyword arguments):
``scatter`` -- how the per-tetra corner forces reach the nodes:
  "tiled" (default) deterministic two-level gather (per-tile sums in LDS, then per-node sums);
  "csr" deterministic gather of per-tetra corner records through the node CSR;
  "atomic" float atomics in arbitrary order, as the reference's atomicAdd.
  The physics is the same; only the floating-point summation order differs.
``tile_k`` tetra per tile for "tiled" (128, 256 or 512);
``stream_nt`` cache policy of the tiled force pass's per-tetra streams (1 nontemporal, 0 default,
  None: nontemporal once they exceed the 256 MB memory-side cache; ``autotune`` times both);
``shape_from_ref`` the tiled force pass forms each tetra's Dm^-1 and V0 from the FP32 reference
  positions instead of streaming them (40 B per tetra less; the same shape to FP32 rounding);
``fused`` and ``use_graph`` switch the fused step and the graph replay off.
"""

from __future__ import annotations

import json
import os
import time

import numpy as np

from nofluidx3d_amd import _lib, openCL
from nofluidx3d_amd.autotune import default_stream_nt
from nofluidx3d_amd.interactions import (FiniteStrainViscoplastic, LegacyElastic, LinearElastic,
                                        LinearElasticDeviatoric, LinearViscoelastic, Plane, PoroNeoHookean,
                                        SecondOrderNeoHookean, Sphere, Substrate, Tetra, VelocityVerlet)
from nofluidx3d_amd.interactions.LegacyContacts import (AdhesivePlane, AdhesivePlaneSurfaceIntegral, CalcCOM,
                                                        COMForce, HardAdhesivePlaneSurfaceIntegral, OverDamped,
                                                        PointZeroForce, RoundTip, SphereIntegral, TipIntegral,
                                                        VelocityVerlet2)
from nofluidx3d_amd.openCL import getCommandQueue, initializeOpenCLObjects
from nofluidx3d_amd.TetraCell import TetraCell
from nofluidx3d_amd.util.Bridge import Bridge
from nofluidx3d_amd.util.IO import writeVTK


def _law(fn, start, speed, move_steps):
    fn.law = (start, speed, move_steps)
    return fn


def tile_ranges(tile_nodes, count, node_off, max_nodes):
    """Per tile-node (ntiles x max_nodes, padded), the range [node_off[n], node_off[n+1]) of its node
    n's partials, two u32 each (nfx_plan.tile_range: the node-fused step reads it with the tile
    metadata); zero for the padding past a tile's count."""
    ntiles = len(count)
    tn = np.asarray(tile_nodes).reshape(ntiles, max_nodes).astype(np.int64)
    valid = np.arange(max_nodes)[None, :] < np.asarray(count)[:, None]
    out = np.zeros((ntiles, max_nodes, 2), np.uint32)
    out[..., 0] = np.where(valid, node_off[tn], 0)
    out[..., 1] = np.where(valid, node_off[tn + 1], 0)
    return out.reshape(-1)


class Simulation:
    def __init__(self, con, scatter="tiled", fused=True, use_graph=True, mesh=None, verbose=True, tile_k=512,
                 stream_nt=None, shape_from_ref=True, fuse_node_pass=None):
        if isinstance(con, dict):
            self.parameters = dict(con)
        else:
            with open(con) as configfile:
                self.parameters = json.load(configfile)
        if scatter not in ("atomic", "csr", "tiled"):
            raise ValueError(f"scatter must be 'tiled', 'csr' or 'atomic', not {scatter!r}")
        self.scatter = scatter
        self.tile_k = tile_k
        self.stream_nt = stream_nt
        self.shape_from_ref = shape_from_ref
        # tiled MR/SVK/SVKK/viscoplastic steps: each node update rides in the next step's force pass
        # (one launch per step instead of two; needs every node to be a corner of some tetra).  None:
        # by mesh size (autotune.default_node_fused), True/False: forced; the autotuner times both
        self.fuse_node_pass = fuse_node_pass
        self.fused = fused
        self.use_graph = use_graph
        self.verbose = verbose

        initializeOpenCLObjects(self.parameters["gpu"])

        # Calculate all conversion constants (Simulation.py:42-56)
        rho = 1.0
        self.rho0 = self.parameters["RhoSI"] / rho
        self.L0 = self.parameters["CELL"]["RadiusSI"] / self.parameters["CELL"]["RadiusSIM"]

        E = 1e-2
        self.p0 = self.parameters["CELL"]["YoungsModulusSI"] / E

        self.V0 = np.sqrt(self.p0 / self.rho0)
        self.T0 = self.L0 / self.V0

        self._print(f"V0={self.V0}")
        self._print(f"T0={self.T0}")

        self.Nsteps = int(round((self.parameters["SimulationTimeSI"]) / self.T0))
        self.VTKInterval = int(round(self.Nsteps / self.parameters["VTKFramesTotal"]))
        self.vtkfilespath = self.parameters["OutputDir"] + "vtkfiles/frame"
        os.makedirs(self.parameters["OutputDir"] + "vtkfiles", exist_ok=True)

        self.cell = TetraCell(
            mesh if mesh is not None else self.parameters["CELL"]["InitVTK"],
            radius=self.parameters["CELL"]["RadiusSIM"],
            recenter=True,
        )

        self.numPoints = self.cell.points[:, 0].size
        self.numTetra = self.cell.tetras[:, 0].size
        # the kernels index node arrays with these ids without a bounds check
        if self.numTetra == 0 or self.cell.tetras.min() < 0 or self.cell.tetras.max() >= self.numPoints:
            raise ValueError("tetra connectivity references nodes outside the mesh")

        self._print(f"Particle mesh consists of {self.numPoints} vertices and {self.numTetra} tetrahedrons")
        self.interactions = []
        self.time = 0

        # device buffers shared by several kernels (Simulation.py:82-90), FP32 state on the device.
        # Positions have two forms: ``points`` holds absolute positions (the reference's buffer,
        # which the per-plugin kernels advance) and ``displacement`` holds u = x - X0 from the
        # reference positions ``refPoints`` (the fused step's state, which keeps FP32 rounding
        # relative to the displacement rather than to |x|).  Whichever was advanced last is
        # current; the other is rebuilt from it on demand.
        self.device = openCL.getContext().torch_device
        self.points = _Positions(self, self.cell.points.reshape(3 * self.numPoints, order="F"))
        self.refPoints = self.points.buf.clone()
        self.displacement = self.points.buf.new_zeros(3 * self.numPoints)
        self._pos_current = {"points", "disp"}
        self.tetras = Bridge(self.cell.tetras.reshape(4 * self.numTetra, order="F"), datatype=np.uint32)
        self.force = Bridge(np.zeros(self.numPoints * 3), datatype=np.float64)
        self.vonMises = Bridge(np.zeros(self.numTetra), datatype=np.float64)
        self.pressure = Bridge(np.zeros(self.numTetra), datatype=np.float64)

        self.recordedQuantities = []

        def currentTime():
            return self.time

        self.recordedQuantities.append(currentTime)
        self._plan = None
        self._plan_key = None
        self._csr = None
        self._tiles = None
        self._tip = None
        self._vp_active = None  # per-tile plastic-activity flags of the tiled viscoplastic step
        self._vp_flags_stale = True
        self.use_plastic_flags = True
        self._node_b = None  # second (disp, vel) buffers of the node-fused step

    # ---------------------------------------------------------------------------------------------
    # the two position forms
    # ---------------------------------------------------------------------------------------------
    def _use_points(self):
        """make the absolute positions current (before per-plugin steps or a read of ``points``)"""
        if "points" not in self._pos_current:
            _lib.call("nfx_points_from_disp", openCL.handle(), self.refPoints.data_ptr(), self.displacement.data_ptr(),
                      self.numPoints, self.points.ptr)
            self._pos_current.add("points")

    def _use_disp(self):
        """make the displacements current (before fused steps)"""
        if "disp" not in self._pos_current:
            _lib.call("nfx_disp_from_points", openCL.handle(), self.refPoints.data_ptr(), self.points.ptr,
                      self.numPoints, self.displacement.data_ptr())
            self._pos_current.add("disp")

    def _print(self, *a, **k):
        if self.verbose:
            print(*a, **k)

    # ---------------------------------------------------------------------------------------------
    # setters (Simulation.py:102-230)
    # ---------------------------------------------------------------------------------------------
    def register(self, interaction):
        interaction.register(self)
        self.interactions.append(interaction)
        self._plan = None

    def setInteractionSubstrate(self):
        # the reference places the wall at -(RadiusSIM + InitialDistance) (Simulation.py:107); an
        # optional "SubstrateDistance" key sets the wall offset independently of the tip start
        dist = self.parameters.get("SubstrateDistance", self.parameters["InitialDistance"])
        wallPosition = -(self.parameters["CELL"]["RadiusSIM"] + dist)
        forceConst = 0.1
        self.register(Substrate(wallPosition, forceConst))

    def _indenter_law(self, start):
        speed = self.parameters["VelocitySI"] / self.V0
        move = self.parameters["MoveTimeSI"] / self.T0

        def pos(time):
            if time <= move:
                return start - speed * time
            return start - speed * move

        return _law(pos, start, speed, move)

    def setInteractionPlane(self, record=True):
        wallFunc = self._indenter_law(self.parameters["CELL"]["RadiusSIM"] + self.parameters["InitialDistance"])
        forceConst = 0.1
        self.register(Plane(wallFunc, forceConst))

        def indentationSI():
            return (wallFunc(0) - wallFunc(self.time)) * self.L0

        def forceSI():
            f, dis = self.tip_force(_lib.INDENTER_PLANE, forceConst, 0.0, wallFunc(self.time))
            self._print(f"Max dis: {dis}")
            return f * (self.p0 * self.L0**2)

        if record:
            self.recordedQuantities.append(indentationSI)
            self.recordedQuantities.append(forceSI)

    def setInteractionSphere(self, record=True, isSpherical=True):
        radius = self.parameters["SphereRadiusSI"] / self.L0
        if isSpherical:
            sphereStartingPos = self.parameters["CELL"]["RadiusSIM"] + radius + self.parameters["InitialDistance"]
        else:
            sphereStartingPos = self.cell.mesh.points[3][1] + radius + self.parameters["InitialDistance"]
        self._print(f"DEBUG: sphereStartingPos = {sphereStartingPos}, lowest point = {sphereStartingPos - radius}")
        sphereFunc = self._indenter_law(sphereStartingPos)
        forceConst = 0.1
        self.register(Sphere(radius, sphereFunc, forceConst))

        def indentationSI():
            return (sphereFunc(0) - sphereFunc(self.time)) * self.L0

        def forceSI():
            f, dis = self.tip_force(_lib.INDENTER_SPHERE, forceConst, radius, sphereFunc(self.time))
            self._print(f"Max dis: {dis}")
            return f * (self.p0 * self.L0**2)

        if record:
            self.recordedQuantities.append(indentationSI)
            self.recordedQuantities.append(forceSI)

    def setInteractionTetra(self):
        youngsModulusSI = self.parameters["CELL"]["YoungsModulusSI"]
        poissonRatio = self.parameters["CELL"]["PoissonRatio"]
        mooneyRivlinRatio = self.parameters["CELL"].get("MooneyRivlinRatio", 1)
        youngsModulus = youngsModulusSI / self.p0
        model = Tetra.materialModel[self.parameters["CELL"]["model"]]
        self.register(Tetra(self.cell, youngsModulus, poissonRatio, mooneyRivlinRatio, model))

    def setInteractionFiniteStrainViscoplastic(self, compactState=False):
        """Legacy parameter mapping of SimulationLegacy.py:842-882 (ReynoldScaling and YoungsScaling
        default to 1 when absent).  Config-2 style keys ``ViscositySI``/``YieldStressSI`` in CELL
        select the Maxwell-type mapping m = 1, n = 0, Y = 3*eta*(1/s), rate = T0 [1/s per step],
        yield offset = YieldStressSI.  ``compactState``: the fused tiled step keeps the symmetric
        plastic deviation only (same forces; no Fp, no per-plugin step; see FiniteStrainViscoplastic)."""
        c = self.parameters["CELL"]
        rs = self.parameters.get("ReynoldScaling", 1.0)
        ys = c.get("YoungsScaling", 1.0)
        youngsModulus = c["YoungsModulusSI"] / rs / ys / self.p0
        nu = c["PoissonRatio"]
        ratio = c.get("MooneyRivlinRatio", 1.0)
        shear = youngsModulus / (2 * (1 + nu))
        bulk = youngsModulus / (3 * (1 - 2 * nu))
        if "ViscositySI" in c:
            flowExponent, hardeningThreshhold, hardeningExponent = 1.0, 1.0, 0.0
            flowRate = self.T0
            yieldStress = 3.0 * c["ViscositySI"] * 1.0 / self.p0
            yieldOffset = c.get("YieldStressSI", 0.0) / self.p0
        else:
            flowExponent = c["ViscoplasticFlowExponent"]
            hardeningThreshhold = c["HardeningThreshhold"]
            hardeningExponent = c["HardeningExponent"]
            flowRate = self.T0
            yieldStress = c["ViscoplasticTimeConstantSI"] ** (1.0 / flowExponent) / ys
            yieldOffset = 0.0
        self.register(FiniteStrainViscoplastic(self.cell, ratio * shear, (1 - ratio) * shear, bulk, flowExponent,
                                               yieldStress, flowRate, hardeningThreshhold, hardeningExponent,
                                               yieldOffset, compactState=compactState))

    # legacy stress laws (SimulationLegacy.py:696-840); ReynoldScaling and YoungsScaling default to 1
    def _legacy_youngs(self):
        c = self.parameters["CELL"]
        return c["YoungsModulusSI"] / self.parameters.get("ReynoldScaling", 1.0) / c.get("YoungsScaling", 1.0) / self.p0

    def setInteractionLinearElastic(self):
        self.register(LinearElastic(self.cell, self._legacy_youngs(), self.parameters["CELL"]["PoissonRatio"]))

    def setInteractionLinearPoroElastic(self):
        c = self.parameters["CELL"]
        E, nu = self._legacy_youngs(), c["PoissonRatio"]
        shear = E / (2 * (1 + nu))
        bulk = (1 - c["PoroElasticFactor"]) * E / (3 * (1 - 2 * nu))
        self.register(LinearElasticDeviatoric(self.cell, shear, bulk))

    def setInteractionSecondOrderNeoHookean(self):
        c = self.parameters["CELL"]
        E, nu = self._legacy_youngs(), c["PoissonRatio"]
        shear = E / (2 * (1 + nu))
        self.register(SecondOrderNeoHookean(self.cell, shear, shear * c["SecondOrderNeoHookeanFactor"],
                                            E / (3 * (1 - 2 * nu))))

    def setInteractionLinearViscoelastic(self):
        """Prony series from CELL["ViscoelasticWeights"] and the relaxation times in seconds
        (CELL["ViscoelasticRelaxationTimesSI"], converted to steps with T0) or in steps
        (CELL["ViscoelasticRelaxationTimes"])."""
        c = self.parameters["CELL"]
        if "ViscoelasticRelaxationTimesSI" in c:
            taus = [t / self.T0 for t in c["ViscoelasticRelaxationTimesSI"]]
        else:
            taus = list(c["ViscoelasticRelaxationTimes"])
        inter = LinearViscoelastic(self.cell, self._legacy_youngs(), c["PoissonRatio"], list(c["ViscoelasticWeights"]),
                                   taus)
        self.register(inter)
        return inter

    def setInteractionPoroNeoHookean(self):
        c = self.parameters["CELL"]
        shear = self._legacy_youngs() / (2 * (1 + 0.5))
        self.register(PoroNeoHookean(self.cell, shear, c["PorousVolumeFraction"]))

    # legacy contacts and integrators (SimulationLegacy.py:139-330, 550-617, 884-940)
    def _tip_start(self, radius, isSpherical):
        if isSpherical:
            return self.parameters["CELL"]["RadiusSIM"] + radius + self.parameters["InitialDistance"]
        return self.cell.points[3][1] + radius + self.parameters["InitialDistance"]

    def _legacy_tip(self, cls, record, isSpherical, **extra):
        radius = self.parameters["SphereRadiusSI"] / self.L0
        fn = self._indenter_law(self._tip_start(radius, isSpherical))
        forceConst = self.parameters["PotentialForceConst"]
        inter = cls(radius=radius, sphereFunc=fn, forceConst=forceConst, **extra)
        self.register(inter)
        if record:
            scale = self.p0 * self.L0**2 * self.parameters.get("ReynoldScaling", 1.0) * \
                self.parameters["CELL"].get("YoungsScaling", 1.0)

            def indentationSI():
                return (fn(0) - fn(self.time)) * self.L0

            def forceSI():
                if cls is RoundTip:  # SimulationLegacy.py:185-200: exp law summed over the nodes
                    x = self.positions()
                    c = np.array([0, fn(self.time), 0])
                    L = np.linalg.norm(x - c, axis=-1)
                    n = (x - c) / L[:, None]
                    return -np.sum(np.exp(-forceConst * (L - radius)) * n[:, 1]) * scale
                # SimulationLegacy.py:248-271: the same integral into a separate force buffer
                f = self._record_force(inter)
                return -np.sum(f[:, 1]) * scale

            self.recordedQuantities.append(indentationSI)
            self.recordedQuantities.append(forceSI)
        return inter

    def _record_force(self, inter):
        import torch

        self._use_points()
        saved = self.force.buf.clone()
        self.force.buf.zero_()
        torch.cuda.synchronize()
        inter.beforeTimeStep(self.time)
        inter.enqueue()
        getCommandQueue().finish()
        f = self.force.buf.cpu().numpy().astype(np.float64).reshape(3, -1).T.copy()
        self.force.buf.copy_(saved)
        torch.cuda.synchronize()
        return f

    def setInteractionRoundTip(self, record=True, isSpherical=True):
        return self._legacy_tip(RoundTip, record, isSpherical)

    def setInteractionSphereIntegral(self, record=True, isSpherical=True):
        return self._legacy_tip(SphereIntegral, record, isSpherical)

    def setInteractionTipIntegral(self, record=True, isSpherical=True):
        halfAngle = self.parameters["TipHalfAngleDeg"] / 360.0 * 2 * np.pi
        return self._legacy_tip(TipIntegral, record, isSpherical, halfAngle=halfAngle,
                                spherePosX=self.cell.points[3][0], spherePosZ=self.cell.points[3][2])

    def _adhesion_wall(self, isSpherical):
        if isSpherical:
            return -self.parameters["CELL"]["RadiusSIM"] - self.parameters["AdhesivePlaneDistance"]
        return self.cell.points[4][1] - self.parameters["AdhesivePlaneDistance"]

    def _adhesion_const(self):
        """AdhesionConstant in simulation units if given, else AdhesionConstantSI over the energy
        unit p0 L0^3 of this unit system (the legacy code divides by its own energy unit E0)."""
        if "AdhesionConstant" in self.parameters:
            return self.parameters["AdhesionConstant"]
        return self.parameters["AdhesionConstantSI"] / (self.p0 * self.L0**3)

    def _surface_triangles(self):
        cd = self.cell.mesh.cells_dict
        if 5 in cd:
            return cd[5]
        from nofluidx3d_amd.interactions.LegacyContacts import surface_triangles

        return surface_triangles(self.cell.tetras)

    def setInteractionAdhesivePlane(self, isSpherical=True):
        wall = self._adhesion_wall(isSpherical)
        on = np.isin(np.arange(self.numPoints), np.unique(self._surface_triangles())).astype(np.int32)
        self.register(AdhesivePlane(lambda t: wall, self._adhesion_const(), self.parameters["AdhesivePlaneDistance"],
                                    on))

    def setInteractionAdhesivePlaneSurfaceIntegral(self, isSpherical=True):
        wall = self._adhesion_wall(isSpherical)
        self.register(AdhesivePlaneSurfaceIntegral(lambda t: wall, self._adhesion_const(),
                                                   self.parameters["AdhesivePlaneDistance"],
                                                   self._surface_triangles()))

    def setInteractionHardAdhesivePlaneNonDimensional(self, isSpherical=True):
        """SimulationLegacy.py:657-692: the wall as for the adhesive planes; the adhesion constant from
        the non-dimensional CELL-independent ``AdhesionConstant`` scaled by E R / (p0 L0) (and the
        legacy ReynoldScaling / YoungsScaling, 1 when absent); k = ``PotentialForceConst``."""
        wall = self._adhesion_wall(isSpherical)
        c = self.parameters["CELL"]
        adhesion = (self.parameters["AdhesionConstant"] * c["YoungsModulusSI"] * c["RadiusSI"] / self.p0 / self.L0 /
                    self.parameters.get("ReynoldScaling", 1.0) / c.get("YoungsScaling", 1.0))
        inter = HardAdhesivePlaneSurfaceIntegral(lambda t: wall, adhesion, self.parameters["PotentialForceConst"],
                                                 self._surface_triangles())
        self.register(inter)
        return inter

    def setInteractionCOMForce(self):
        calc = CalcCOM()
        self.register(calc)
        self.register(COMForce(self.parameters["COMSpringConst"], calc))
        self.com = calc

    def setInteractionPointZeroForce(self):
        self.register(PointZeroForce(self.parameters["COMSpringConst"]))

    # integrators: register last
    def setInteractionVelocityVerlet2(self, damping=0.1):
        from nofluidx3d_amd.interactions.LegacyContacts import lumped_masses

        self.register(VelocityVerlet2(lumped_masses(self.cell.tetras, self.cell.volumes, self.numPoints), damping))

    def setInteractionOverDamped(self, point3fix):
        self.register(OverDamped(point3fix))

    # Velocity Verlet Interaction, always add time integration interactions as last element
    def setInteractionVelocityVerlet(self, fixTopBottom=False):
        self.register(VelocityVerlet(self.numPoints, fixTopBottom))

    # ---------------------------------------------------------------------------------------------
    # stepping
    # ---------------------------------------------------------------------------------------------
    def node_csr(self):
        """Deterministic-gather buffers: CSR offsets/slots of node->corner incidences and the
        per-corner force scratch (12 floats per tetra)."""
        if self._csr is None:
            import ctypes

            import torch

            tets = np.ascontiguousarray(self.tetras.arrayC, dtype=np.uint32)
            off = np.zeros(self.numPoints + 1, dtype=np.uint32)
            slots = np.zeros(4 * self.numTetra, dtype=np.uint32)
            _lib.call("nfx_build_node_csr_host", tets.ctypes.data_as(ctypes.c_void_p), self.numPoints, self.numTetra,
                      off.ctypes.data_as(ctypes.c_void_p), slots.ctypes.data_as(ctypes.c_void_p))
            dev = self.device
            self._csr = (
                torch.from_numpy(off.view(np.int32)).to(dev),
                torch.from_numpy(slots.view(np.int32)).to(dev),
                torch.empty(12 * self.numTetra, dtype=torch.float32, device=dev),
            )
            torch.cuda.synchronize(dev)
        return self._csr

    def node_tiles(self):
        """Tiled-gather arrays (see nfx_build_tiles_host) on the device."""
        if self._tiles is None:
            import ctypes

            import torch

            tets = np.ascontiguousarray(self.tetras.arrayC, dtype=np.uint32)
            K = int(self.tile_k)
            ntiles = (self.numTetra + K - 1) // K
            count = np.zeros(ntiles, dtype=np.uint32)
            m, mx = ctypes.c_int64(), ctypes.c_int32()
            vp = ctypes.c_void_p
            _lib.call("nfx_build_tiles_host", tets.ctypes.data_as(vp), self.numPoints, self.numTetra, K,
                      ctypes.byref(mx), ctypes.byref(m), count.ctypes.data_as(vp), None, None, None, None, None)
            MX, M = mx.value, m.value
            if MX > min(K, _lib.TILE_MAX_NODES):  # not tileable at this size: the plan falls back to CSR
                self._tiles = {"tile_k": K, "max_nodes": MX, "m_total": M}
                return self._tiles
            tile_nodes = np.zeros(ntiles * MX, np.uint32)
            tile_lend = np.zeros(ntiles * MX, np.uint16)
            tile_corners = np.zeros(4 * K * ntiles, np.uint16)
            tile_dst = np.zeros(ntiles * MX, np.uint32)
            node_off = np.zeros(self.numPoints + 1, np.uint32)
            _lib.call("nfx_build_tiles_host", tets.ctypes.data_as(vp), self.numPoints, self.numTetra, K,
                      ctypes.byref(mx), ctypes.byref(m), count.ctypes.data_as(vp), tile_nodes.ctypes.data_as(vp),
                      tile_lend.ctypes.data_as(vp), tile_corners.ctypes.data_as(vp), tile_dst.ctypes.data_as(vp),
                      node_off.ctypes.data_as(vp))
            dev = self.device
            to = lambda a: torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a.view(np.int16)).to(dev)
            tile_range = tile_ranges(tile_nodes, count, node_off, MX)
            self._tiles = {
                "tile_k": K,
                "max_nodes": MX,
                "m_total": M,
                "tile_count": to(count),
                "tile_nodes": to(tile_nodes),
                "tile_lend": to(tile_lend),
                "tile_corners": to(tile_corners),
                "tile_dst": to(tile_dst),
                "node_off": to(node_off),
                "tile_range": to(tile_range),
                "partials": torch.zeros(3 * M, dtype=torch.float32, device=dev),
                # every node is a corner of some tetra (has a partial): the node-fused step updates
                # nodes only through the tiles that touch them
                "every_node": bool(np.all(node_off[1:] > node_off[:-1])),
            }
            torch.cuda.synchronize(dev)
        return self._tiles

    def plan(self):
        """The fused-step description of the registered plugin chain, or None if it cannot fuse.
        Plugin state is brought into the form the fused step advances (callers step right after)."""
        p = self._build_plan()
        if p is not None:
            self._use_disp()
            for inter in self.interactions:
                inter.state_form(fused=True)
            if p.vp_active and self._vp_flags_stale:
                _lib.call("nfx_vp_tile_flags", openCL.handle(), p)
                self._vp_flags_stale = False
        return p

    def plastic_tile_fraction(self):
        """Fraction of the tiled viscoplastic step's tiles whose tetra have flowed (None without flags)."""
        p = self._plan
        if self._vp_active is None or p is None or not p.vp_active:
            return None
        openCL.getCommandQueue().finish()
        return float((self._vp_active.cpu().numpy() != 0).mean())

    def plastic_state_changed(self):
        """The viscoplastic state was written outside the fused tiled step (a conversion, a restore,
        a caller's own write): the per-tile activity flags are recomputed before the next fused step."""
        self._vp_flags_stale = True

    def sync_forms(self):
        """Bring every state representation up to date (before a snapshot of the state)."""
        self._use_points()
        self._use_disp()
        for inter in self.interactions:
            inter.sync_forms()

    def forms_restored(self):
        self._pos_current = {"points", "disp"}
        self._vp_flags_stale = True
        for inter in self.interactions:
            inter.forms_restored()

    def _build_plan(self):
        if not self.fused:
            return None
        key = tuple(id(i) for i in self.interactions) + (self.scatter, self.node_fused_wanted())
        if self._plan is not None and self._plan_key == key:
            return self._plan
        p = _lib.Plan()
        p.N, p.T = self.numPoints, self.numTetra
        p.tetras, p.force = self.tetras.ptr, self.force.ptr
        p.ref, p.disp = self.refPoints.data_ptr(), self.displacement.data_ptr()
        ok = len(self.interactions) >= 2 and isinstance(self.interactions[-1], VelocityVerlet)
        for inter in self.interactions:
            ok = ok and inter.contribute(p)
        if not ok or not p.dminv or not p.vel:
            self._plan, self._plan_key = None, key
            return None
        if self.scatter == "csr":
            off, slots, corners = self.node_csr()
            p.scatter = _lib.SCATTER_CSR
            p.csr_offsets, p.csr_slots, p.corners = off.data_ptr(), slots.data_ptr(), corners.data_ptr()
        elif self.scatter == "tiled" and self.node_tiles()["max_nodes"] > min(self.node_tiles()["tile_k"],
                                                                               _lib.TILE_MAX_NODES):
            # a tile touching more distinct nodes than it has threads (or than its 8-bit tile-node
            # index allows): use the CSR gather instead
            self.scatter = "csr"
            return self._build_plan()
        elif self.scatter == "tiled":
            t = self.node_tiles()
            p.scatter = _lib.SCATTER_TILED
            p.tile_k, p.tile_max_nodes = t["tile_k"], t["max_nodes"]
            p.stream_nt = default_stream_nt(self.numTetra) if self.stream_nt is None else int(self.stream_nt)
            p.shape_from_ref = 1 if self.shape_from_ref else 0
            if p.model == _lib.NFX_VISCOPLASTIC and self.use_plastic_flags:
                import torch

                ntiles = -(-self.numTetra // t["tile_k"])
                if self._vp_active is None or self._vp_active.numel() != ntiles:
                    self._vp_active = torch.ones(ntiles, dtype=torch.int32, device=self.device)
                p.vp_active = self._vp_active.data_ptr()
                self._vp_flags_stale = True
            for k in ("tile_count", "tile_nodes", "tile_lend", "tile_corners", "tile_dst", "node_off", "partials"):
                setattr(p, k, t[k].data_ptr())
            if self.node_fused_wanted() and t["every_node"] and p.model in (_lib.NFX_SVKK, _lib.NFX_SVK, _lib.NFX_MR,
                                                                        _lib.NFX_VISCOPLASTIC):
                import torch

                if self._node_b is None:
                    self._node_b = torch.zeros(2, 3 * self.numPoints, dtype=torch.float32, device=self.device)
                if "partials_b" not in t:
                    t["partials_b"] = torch.zeros_like(t["partials"])
                p.disp_b, p.vel_b = self._node_b[0].data_ptr(), self._node_b[1].data_ptr()
                p.partials_b = t["partials_b"].data_ptr()
                p.tile_range = t["tile_range"].data_ptr()
        else:
            p.scatter = _lib.SCATTER_ATOMIC
        self._plan, self._plan_key = p, key
        return p

    def node_fused_wanted(self):
        """whether the fused tiled step should take the node-fused form (fuse_node_pass, or by mesh
        size when it is None)"""
        if self.fuse_node_pass is None:
            from nofluidx3d_amd.autotune import default_node_fused

            return default_node_fused(self.numTetra)
        return bool(self.fuse_node_pass)

    def advance(self, nsteps, tip_hist=None):
        """Advance ``nsteps`` steps (fused when possible).  ``tip_hist``: optional float64 device
        tensor receiving the indenter force (simulation units) of each step."""
        p = self.plan()
        if p is None:
            if tip_hist is not None:
                raise _lib.NfxError("per-step tip history needs the fused step")
            for _ in range(nsteps):
                self.timeStep()
            return
        _lib.call("nfx_run_steps", openCL.handle(), p, self.time, nsteps,
                  None if tip_hist is None else tip_hist.data_ptr(), 1 if self.use_graph else 0)
        self._pos_current = {"disp"}
        self.time += nsteps

    def timeStep(self):
        p = self.plan()
        if p is not None:
            _lib.call("nfx_run_steps", openCL.handle(), p, self.time, 1, None, 0)
            self._pos_current = {"disp"}
            self.time += 1
            return
        self._use_points()
        for interaction in self.interactions:
            interaction.beforeTimeStep(self.time)
        for interaction in self.interactions:
            interaction.enqueue()
        self._pos_current = {"points"}
        self.time += 1

    def tip_force(self, indenter, k, radius, pos):
        """FP64 device reduction of the indenter force at the current positions (simulation
        units) and the maximum penetration."""
        import torch

        if self._tip is None:
            self._tip = torch.zeros(2, dtype=torch.float64, device=self.device)
        if "disp" in self._pos_current:
            x, ref = self.displacement.data_ptr(), self.refPoints.data_ptr()
        else:
            x, ref = self.points.ptr, None
        _lib.call("nfx_tip_force", openCL.handle(), x, ref, self.numPoints, indenter, float(k), float(radius),
                  float(pos), self._tip.data_ptr())
        getCommandQueue().finish()
        out = self._tip.cpu().numpy()
        return float(out[0]), float(out[1])

    def save_state(self, path):
        """Checkpoint to an ``.npz`` file: the step counter and every state tensor (node positions,
        displacements, forces, stress outputs, each plugin's state).  ``load_state`` restores it into
        a Simulation built the same way (same config, mesh and plugin chain)."""
        import torch

        self.sync_forms()
        openCL.getCommandQueue().finish()
        torch.cuda.synchronize(self.device)
        arrays = {f"t{i}": t.detach().cpu().numpy() for i, t in enumerate(self.state_tensors())}
        np.savez(path, time=np.int64(self.time), count=np.int64(len(arrays)), **arrays)

    def load_state(self, path):
        import torch

        d = np.load(path)
        tensors = self.state_tensors()
        if int(d["count"]) != len(tensors):
            raise ValueError(f"{path}: {int(d['count'])} state tensors, this simulation has {len(tensors)}")
        for i, t in enumerate(tensors):
            a = d[f"t{i}"]
            if a.shape != tuple(t.shape) or a.dtype != torch.empty(0, dtype=t.dtype).numpy().dtype:
                raise ValueError(f"{path}: tensor {i} is {a.dtype}{a.shape}, expected {tuple(t.shape)}")
        openCL.getCommandQueue().finish()
        for i, t in enumerate(tensors):
            t.copy_(torch.from_numpy(d[f"t{i}"]))
        torch.cuda.synchronize(self.device)
        self.forms_restored()
        self.time = int(d["time"])

    def state_tensors(self):
        """Device tensors that carry the simulation state (positions, force, plugin state); call
        ``sync_forms`` before taking them and ``forms_restored`` after writing them back."""
        out = [self.points.buf, self.displacement, self.force.buf, self.vonMises.buf, self.pressure.buf]
        for inter in self.interactions:
            out.extend(inter.state_buffers())
        return out

    def positions(self) -> np.ndarray:
        """Current node positions (N, 3) float64.  From the displacement form they are X0 + u
        summed in float64 (both are FP32 values), which keeps the displacement's precision."""
        if "disp" in self._pos_current:
            getCommandQueue().finish()
            ref = self.refPoints.cpu().numpy().astype(np.float64)
            u = self.displacement.cpu().numpy().astype(np.float64)
            return (ref + u).reshape(3, self.numPoints).T.copy()
        return self.points.readFromGPU().reshape(3, self.numPoints).T.copy()

    def _record(self):
        return [quantity() for quantity in self.recordedQuantities]

    def run(self, writeStress=True, writePressure=True, writeVolumeChange=False):
        """The reference's run loop (``Simulation.py:239-290``): a frame and a row of recorded
        quantities at t = 0 and after every ``VTKInterval`` steps, ``VTKFramesTotal`` times;
        ``simdata.dat`` is rewritten with all rows after each frame."""
        self._print("\n\n")
        self._print("+------------------------------------+")
        self._print("|         Starting Simulation        |")
        self._print("+------------------------------------+")
        self._print("\n\n")
        self._print(f"VTKInterval: {self.VTKInterval}")
        for inter in self.interactions:
            if isinstance(inter, (FiniteStrainViscoplastic, LegacyElastic, LinearViscoelastic)) and \
                    (writeStress or writePressure):
                inter.write_stress = True
                self._plan = None
        frames = self.parameters["VTKFramesTotal"]
        simdata = self.parameters["OutputDir"] + "simdata.dat"

        def frame():
            writeVTK(self, writeStress=writeStress, writePressure=writePressure, writeVolumeChange=writeVolumeChange)
            rows.append(self._record())
            np.savetxt(simdata, np.array(rows))

        rows = []
        frame()
        step_seconds = []  # wall time of each block of VTKInterval steps
        for done in range(1, frames + 1):
            t_block = time.time()
            self.advance(self.VTKInterval)
            getCommandQueue().finish()
            t_frame = time.time()
            frame()
            step_seconds.append(t_frame - t_block)
            frame_seconds = time.time() - t_frame
            mean_block = sum(step_seconds) / len(step_seconds)
            remaining_min = (frames - done + 1) * (mean_block + frame_seconds) / 60
            self._print(
                f"| Average # of simulation steps per second: {np.round(self.VTKInterval / mean_block, 1)} "
                f"| VTK cycle took {np.round(frame_seconds, 1)} seconds | estimate of time remaining: "
                f"{np.round(remaining_min, 1)} min",
                end="\r",
            )
        self._print("\n\n")
        self._print("+------------------------------------+")
        self._print("|  Simulation finished successfully  |")
        self._print("+------------------------------------+")


class _Positions(Bridge):
    """The ``points`` buffer: reading it (``readFromGPU``) first rebuilds the absolute positions
    when the displacement form was advanced last, and writing it (``writeToGPU``) makes the
    absolute form the current one."""

    def __init__(self, simulation, host):
        super().__init__(host, datatype=np.float64)
        self._simulation = simulation

    def readFromGPU(self):
        self._simulation._use_points()
        return super().readFromGPU()

    def writeToGPU(self):
        super().writeToGPU()
        self._simulation._pos_current = {"points"}
