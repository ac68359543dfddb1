"""Per-tetra hyperelastic force (reference ``interactions/Tetra.py:23-92``, kernel ``tetra.cl``).

``createBuffers`` uploads the reference cell's FP64 edge vectors and volumes once and inverts the
reference shape matrices on the device in FP64 (``nfx_precompute_f64``); the step kernel reads the
FP32 inverse and volume.  The model flags are the reference's ``#define``s (``Tetra.py:66-79``);
NEO_HOOKEAN and MOONEY_RIVLIN share one branch.

Node scatter: ``simulation.scatter == "atomic"`` adds the corner forces with float atomics (the
reference's ``atomicAdd``); ``"csr"`` (and ``"tiled"``, whose two-level gather exists only in the
fused step) writes one corner record per tetra and gathers them per node in a fixed order, which
makes the force sum bitwise reproducible.
"""
from enum import Enum, auto
from pathlib import Path

import numpy as np

from nofluidx3d_amd import _lib, openCL
from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.KernelBuilder import KernelBuilder

kernels = Path(__file__).parents[1] / "kernels"

MODEL_CODE = {}


class Tetra(Interaction):
    class materialModel(Enum):
        SAINT_VENANT_KIRCHHOFF_KELLNBERGER = auto()
        SVKK = SAINT_VENANT_KIRCHHOFF_KELLNBERGER
        SAINT_VENANT_KIRCHHOFF = auto()
        SVK = SAINT_VENANT_KIRCHHOFF
        NEO_HOOKEAN = auto()
        NH = NEO_HOOKEAN
        MOONEY_RIVLIN = auto()
        MR = MOONEY_RIVLIN

    def __init__(
        self,
        referenceCell,
        youngsModulus,
        poissonRatio,
        mooneyRivlinRatio=1,
        model=materialModel.NEO_HOOKEAN,
    ):
        self.referenceCell = referenceCell
        self.youngsModulus = youngsModulus
        self.poissonRatio = poissonRatio
        self.mooneyRivlinRatio = mooneyRivlinRatio
        self.model = model

    def createBuffers(self):
        self.dminvB, self.volumesB = precompute_reference(self.referenceCell, self.simulation)
        self.referenceCell = None

    def build(self):
        self.scope = self.simulation.numTetra
        M = self.materialModel
        KernelBuilder.define(
            def_pointCount=self.simulation.numPoints,
            def_tetraCount=self.simulation.numTetra,
            def_tetraYoungsModulus=self.youngsModulus,
            def_tetraPoissonRatio=self.poissonRatio,
            def_tetraMooneyRivlinRatio=self.mooneyRivlinRatio,
            TETRA_SAINT_VENANT_KIRCHHOFF_KELLNBERGER=self.model == M.SAINT_VENANT_KIRCHHOFF_KELLNBERGER,
            TETRA_SAINT_VENANT_KIRCHHOFF=self.model == M.SAINT_VENANT_KIRCHHOFF,
            TETRA_MOONEY_RIVLIN=self.model == M.MOONEY_RIVLIN or self.model == M.NEO_HOOKEAN,
        )
        self.knl = KernelBuilder.build(kernels / "Interactions" / "tetra.cl", "calcTetraForce")

    def setArgs(self):
        self.knl.set_args(
            self.simulation.force.buf,
            self.simulation.points.buf,
            self.simulation.tetras.buf,
            self.dminvB,
            self.volumesB,
        )

    def model_code(self):
        M = self.materialModel
        if self.model == M.SAINT_VENANT_KIRCHHOFF_KELLNBERGER:
            return _lib.NFX_SVKK
        if self.model == M.SAINT_VENANT_KIRCHHOFF:
            return _lib.NFX_SVK
        return _lib.NFX_MR

    def enqueue(self):
        sim = self.simulation
        if getattr(sim, "scatter", "atomic") in ("csr", "tiled"):
            off, slots, corners = sim.node_csr()
            v = self.knl.values
            mat = _lib.Elastic(self.model_code(), float(v["def_tetraYoungsModulus"]),
                               float(v["def_tetraPoissonRatio"]), float(v["def_tetraMooneyRivlinRatio"]))
            h = openCL.handle()
            _lib.call("nfx_tetra_corner_forces", h, mat, sim.points.ptr, sim.tetras.ptr, self.dminvB.data_ptr(),
                      self.volumesB.data_ptr(), sim.numPoints, sim.numTetra, corners.data_ptr())
            _lib.call("nfx_gather_corners", h, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), sim.numPoints,
                      sim.force.ptr)
        else:
            self.knl()

    def contribute(self, plan) -> bool:
        if plan.dminv:
            return False
        v = self.knl.values
        plan.model = self.model_code()
        plan.E = float(v["def_tetraYoungsModulus"])
        plan.nu = float(v["def_tetraPoissonRatio"])
        plan.ratio = float(v["def_tetraMooneyRivlinRatio"])
        plan.dminv = self.dminvB.data_ptr()
        plan.v0 = self.volumesB.data_ptr()
        return True


def precompute_reference(referenceCell, simulation):
    """Upload FP64 edge vectors (SoA, Tetra.py:53-58) and volumes; return device (Dm^-1, V0) FP32."""
    import torch

    T = simulation.numTetra
    dev = simulation.points.buf.device
    ev = torch.from_numpy(np.ascontiguousarray(referenceCell.edgeVectors.reshape(9 * T, order="F"),
                                               dtype=np.float64)).to(dev)
    vol = torch.from_numpy(np.ascontiguousarray(referenceCell.volumes, dtype=np.float64)).to(dev)
    dminv = torch.empty(9 * T, dtype=torch.float32, device=dev)
    v0 = torch.empty(T, dtype=torch.float32, device=dev)
    torch.cuda.synchronize(dev)
    _lib.call("nfx_precompute_f64", openCL.handle(), ev.data_ptr(), vol.data_ptr(), T, dminv.data_ptr(),
              v0.data_ptr())
    openCL.getCommandQueue().finish()
    del ev, vol
    return dminv, v0
