"""Finite-strain viscoplastic material (multiplicative Fe Fp split, power-law flow with hardening).

Restates the legacy ``Interaction_FiniteStrainViscoplastic`` (reference
``kernels/InteractionFiniteStrainViscoplastic.cl:1-220``, wrapper ``Interactions.py:248-349``)
behind the current plugin protocol, with node 0 as the edge base (the convention of
``TetraCell.py:33-38``; the legacy kernel used node 3).  Parameters are uniform over the mesh, as
the legacy wrapper's eight per-tetra arrays are (``Interactions.py:270-277``).  ``yieldOffset``
(default 0 = the reference law) makes the flow rate use max(svm - yieldOffset, 0).

State per tetra: the plastic deformation gradient Fp (9 SoA, starts at I) and the accumulated
plastic strain; ``vonMises`` and ``pressure`` are written when the plugin's ``write_stress`` is on.
The tiled fused step keeps the same state in shape form S = Dm^-1 Fp^-1 (``shapeB``), which
replaces both Dm^-1 and Fp in its force pass; ``state_form`` converts between the two forms when
the stepping path changes, and ``plastic_deformation()`` returns an up-to-date Fp.
"""
from pathlib import Path

import numpy as np

from nofluidx3d_amd import _lib
from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.interactions.Tetra import precompute_reference
from nofluidx3d_amd.KernelBuilder import KernelBuilder

kernels = Path(__file__).parents[1] / "kernels"


class FiniteStrainViscoplastic(Interaction):
    def __init__(
        self,
        referenceCell,
        shearModulus1,
        shearModulus2,
        bulkModulus,
        viscoplasticFlowExponent,
        viscoplasticYieldStress,
        viscoplasticFlowRate,
        hardeningThreshhold,
        hardeningExponent,
        yieldOffset=0.0,
    ):
        self.referenceCell = referenceCell
        self.params = _lib.Viscoplastic(
            float(shearModulus1), float(shearModulus2), float(bulkModulus), float(viscoplasticFlowExponent),
            float(viscoplasticYieldStress), float(viscoplasticFlowRate), float(hardeningThreshhold),
            float(hardeningExponent), float(yieldOffset))
        self.write_stress = False

    def createBuffers(self):
        import torch

        self.dminvB, self.volumesB = precompute_reference(self.referenceCell, self.simulation)
        self.referenceCell = None
        T = self.simulation.numTetra
        dev = self.dminvB.device
        self.accumulatedStrainB = torch.zeros(T, dtype=torch.float32, device=dev)
        # identity in the SoA layout: entry k of tetra t at k*T + t (Interactions.py:279-281)
        fp = np.repeat(np.identity(3).reshape(9), T).astype(np.float32)
        self.plasticDeformationTensorB = torch.from_numpy(fp).to(dev)
        self.shapeB = self.dminvB.clone()  # S = Dm^-1 while Fp = I
        torch.cuda.synchronize(dev)
        self._current = {"fp", "shape"}

    def state_form(self, shape: bool):
        """Make Fp (shape=False) or S (shape=True) current before steps that advance it; after
        them only that form is current."""
        want = "shape" if shape else "fp"
        if want not in self._current:
            self._convert(want)
        self._current = {want}

    def _convert(self, want):
        from nofluidx3d_amd import openCL

        T = self.simulation.numTetra
        if want == "shape":
            _lib.call("nfx_vp_shape_from_fp", openCL.handle(), self.dminvB.data_ptr(),
                      self.plasticDeformationTensorB.data_ptr(), T, self.shapeB.data_ptr())
        else:
            _lib.call("nfx_vp_fp_from_shape", openCL.handle(), self.dminvB.data_ptr(), self.shapeB.data_ptr(), T,
                      self.plasticDeformationTensorB.data_ptr())
        self._current.add(want)

    def sync_forms(self):
        """Bring both forms up to date (no steps in between)."""
        for want in ("fp", "shape"):
            if want not in self._current:
                self._convert(want)

    def forms_restored(self):
        """Both forms were just restored from one snapshot taken after ``sync_forms``."""
        self._current = {"fp", "shape"}

    def plastic_deformation(self):
        """The plastic deformation gradient Fp (9T SoA device tensor), up to date."""
        from nofluidx3d_amd import openCL

        self.sync_forms()
        openCL.getCommandQueue().finish()
        return self.plasticDeformationTensorB

    def build(self):
        self.scope = self.simulation.numTetra
        KernelBuilder.define(INSERT_NUM_POINTS=self.simulation.numPoints, INSERT_NUM_TETRAS=self.simulation.numTetra)
        self.knl = KernelBuilder.build(kernels / "InteractionFiniteStrainViscoplastic.cl",
                                       "Interaction_FiniteStrainViscoplastic")

    def setArgs(self):
        sim = self.simulation
        self.knl.set_args(sim.force.buf, sim.points.buf, sim.tetras.buf, self.dminvB, self.params,
                          self.accumulatedStrainB, self.plasticDeformationTensorB,
                          sim.vonMises.buf if self.write_stress else None,
                          sim.pressure.buf if self.write_stress else None)

    def enqueue(self):
        sim = self.simulation
        self.state_form(shape=False)
        if getattr(sim, "scatter", "atomic") in ("csr", "tiled"):
            from nofluidx3d_amd import openCL

            off, slots, corners = sim.node_csr()
            h = openCL.handle()
            vm = sim.vonMises.ptr if self.write_stress else None
            pr = sim.pressure.ptr if self.write_stress else None
            _lib.call("nfx_viscoplastic_force", h, self.params, sim.points.ptr, sim.tetras.ptr,
                      self.dminvB.data_ptr(), sim.numPoints, sim.numTetra, self.plasticDeformationTensorB.data_ptr(),
                      self.accumulatedStrainB.data_ptr(), vm, pr, None, corners.data_ptr())
            _lib.call("nfx_gather_corners", h, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), sim.numPoints,
                      sim.force.ptr)
        else:
            self.knl()

    def contribute(self, plan) -> bool:
        if plan.dminv:
            return False
        plan.model = _lib.NFX_VISCOPLASTIC
        plan.vp = self.params
        plan.dminv = self.dminvB.data_ptr()
        plan.v0 = self.volumesB.data_ptr()
        plan.Fp = self.plasticDeformationTensorB.data_ptr()
        plan.vp_shape = self.shapeB.data_ptr()
        plan.epsbar = self.accumulatedStrainB.data_ptr()
        if self.write_stress:
            plan.von_mises, plan.pressure = self.simulation.vonMises.ptr, self.simulation.pressure.ptr
        return True
