"""Per-tetra hyperelastic force (reference ``interactions/Tetra.py:23-92``, kernel
``Interactions/tetra.cl:14-182``).

The reference re-inverts every tetra's reference shape matrix Dm = [R1 | R2 | R3] each step from
the edge vectors it uploads.  Here ``createBuffers`` uploads the cell's FP64 edge vectors and
volumes once and ``nfx_precompute_f64`` inverts Dm on the device in FP64; the step kernels read
the FP32 inverse and V0.  The material constants, which the reference bakes into its kernel as
``#define``s, travel as arguments; NEO_HOOKEAN and MOONEY_RIVLIN take the same branch
(``Tetra.py:77-78``: Neo-Hookean is Mooney-Rivlin with ratio 1).

Node scatter of the per-plugin step: ``simulation.scatter == "atomic"`` adds corner forces with
float atomics (the reference's ``atomicAdd``); ``"csr"`` and ``"tiled"`` write a corner record per
tetra and sum them per node in a fixed order (``nfx_gather_corners``), which makes the sum bitwise
reproducible.  The fused step takes the plugin's arrays through ``contribute``.
"""
from enum import Enum, auto

import numpy as np

from nofluidx3d_amd import _lib, openCL
from nofluidx3d_amd.interactions.Interaction import Interaction


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

    def __init__(self, referenceCell, youngsModulus, poissonRatio, mooneyRivlinRatio=1, model=materialModel.NEO_HOOKEAN):
        super().__init__()
        self.referenceCell = referenceCell
        self.youngsModulus = youngsModulus
        self.poissonRatio = poissonRatio
        self.mooneyRivlinRatio = mooneyRivlinRatio
        self.model = model

    def model_code(self):
        """The libnfx model number (NFX_SVKK, NFX_SVK, or NFX_MR for both NH and MR)."""
        codes = {self.materialModel.SAINT_VENANT_KIRCHHOFF_KELLNBERGER: _lib.NFX_SVKK,
                 self.materialModel.SAINT_VENANT_KIRCHHOFF: _lib.NFX_SVK}
        return codes.get(self.model, _lib.NFX_MR)

    def material(self):
        return _lib.Elastic(self.model_code(), float(self.youngsModulus), float(self.poissonRatio),
                            float(self.mooneyRivlinRatio))

    def createBuffers(self):
        self.dminvB, self.volumesB = precompute_reference(self.referenceCell, self.simulation)
        self.referenceCell = None  # the host copy is not needed any more

    def build(self):
        self.scope = self.simulation.numTetra
        self.knl = self._enqueue_force

    def _enqueue_force(self):
        sim, h = self.simulation, openCL.handle()
        geometry = (sim.tetras.ptr, self.dminvB.data_ptr(), self.volumesB.data_ptr(), sim.numPoints, sim.numTetra)
        if getattr(sim, "scatter", "atomic") == "atomic":
            _lib.call("nfx_tetra_force", h, self.material(), sim.points.ptr, *geometry, sim.force.ptr)
            return
        off, slots, corners = sim.node_csr()
        _lib.call("nfx_tetra_corner_forces", h, self.material(), sim.points.ptr, *geometry, corners.data_ptr())
        _lib.call("nfx_gather_corners", h, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), sim.numPoints,
                  sim.force.ptr)

    def contribute(self, plan) -> bool:
        if plan.dminv:  # one force law per fused step
            return False
        plan.model = self.model_code()
        plan.E, plan.nu, plan.ratio = float(self.youngsModulus), float(self.poissonRatio), float(self.mooneyRivlinRatio)
        plan.dminv, plan.v0 = self.dminvB.data_ptr(), self.volumesB.data_ptr()
        return True


def precompute_reference(referenceCell, simulation):
    """(Dm^-1 9T, V0 T) FP32 device tensors from the cell's FP64 edge vectors and volumes, inverted
    on the device in FP64.  The edge vectors travel in the reference's SoA order (``Tetra.py:53-58``:
    entry k of tetra t at k*T + t)."""
    import torch

    T = simulation.numTetra
    dev = simulation.device
    soa = np.asfortranarray(referenceCell.edgeVectors, dtype=np.float64).T.reshape(-1)
    ev = torch.from_numpy(np.ascontiguousarray(soa)).to(dev)
    vol = torch.from_numpy(np.ascontiguousarray(referenceCell.volumes, dtype=np.float64)).to(dev)
    dminv = torch.empty(9 * T, dtype=torch.float32, device=dev)
    v0 = torch.empty(T, dtype=torch.float32, device=dev)
    torch.cuda.synchronize(dev)
    _lib.call("nfx_precompute_f64", openCL.handle(), ev.data_ptr(), vol.data_ptr(), T, dminv.data_ptr(),
              v0.data_ptr())
    openCL.getCommandQueue().finish()
    return dminv, v0
