"""Finite-strain viscoplastic material (multiplicative Fe Fp split, power-law flow with hardening).

Restates the legacy ``Interaction_FiniteStrainViscoplastic`` (reference
``kernels/InteractionFiniteStrainViscoplastic.cl:1-220``, wrapper ``Interactions.py:248-349``)
behind the current plugin protocol, with node 0 as the edge base (the convention of
``TetraCell.py:33-38``; the legacy kernel used node 3).  Parameters are uniform over the mesh, as
the legacy wrapper's eight per-tetra arrays are (``Interactions.py:270-277``).  ``yieldOffset``
(default 0 = the reference law) makes the flow rate use max(svm - yieldOffset, 0).

State per tetra: the accumulated plastic strain, and the plastic deformation in one of two forms:
- ``plasticDeformationTensorB``: Fp itself (9 SoA, I at the start), the reference's buffer, which
  the per-plugin kernel advances;
- ``plasticDeviationB``: Q = Fp^-1 - I (9 SoA, zero at the start), which every fused step
  advances.  In FP32 a small plastic increment added to Fp is rounded against Fp's O(1) entries
  every step; Q carries the same information with its own relative precision.
``state_form`` converts when the stepping path changes; ``plastic_deformation()`` returns an
up-to-date Fp.  ``vonMises`` and ``pressure`` are written when the plugin's ``write_stress`` is on.

``compactState=True`` keeps, for the fused tiled step only, the symmetric deviation
Gd = Fp^-1 Fp^-T - I (6 SoA, ``plasticDeviationSymB``) instead of Fp and Q, and the accumulated plastic
strain only when the flow reads it (flow exponent != 1 or hardening exponent != 0).  The stress, the
flow and the forces depend on Gd alone, so positions and forces are the same to FP32 rounding with 36
B per tetra and step less traffic (the Maxwell-type law: 44 B); Fp is not recoverable from Gd, so
the per-plugin step and ``plastic_deformation()`` are unavailable.
"""
import numpy as np

from nofluidx3d_amd import _lib, openCL
from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.interactions.Tetra import precompute_reference


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
        compactState=False,
    ):
        super().__init__()
        self.referenceCell = referenceCell
        self.params = _lib.Viscoplastic(
            float(shearModulus1), float(shearModulus2), float(bulkModulus), float(viscoplasticFlowExponent),
            float(viscoplasticYieldStress), float(viscoplasticFlowRate), float(hardeningThreshhold),
            float(hardeningExponent), float(yieldOffset))
        self.write_stress = False
        self.compact = bool(compactState)

    def _reads_strain(self):
        return self.params.flow_exponent != 1.0 or self.params.hardening_exponent != 0.0

    def createBuffers(self):
        import torch

        self.dminvB, self.volumesB = precompute_reference(self.referenceCell, self.simulation)
        self.referenceCell = None
        T = self.simulation.numTetra
        dev = self.simulation.device
        self.accumulatedStrainB = torch.zeros(T, dtype=torch.float32, device=dev)
        if self.compact:
            self.plasticDeviationSymB = torch.zeros(6 * T, dtype=torch.float32, device=dev)
            torch.cuda.synchronize(dev)
            self._current = {"gd"}
            return
        eye_soa = np.repeat(np.eye(3, dtype=np.float32).reshape(9), T)  # entry k of tetra t at k*T + t
        self.plasticDeformationTensorB = torch.from_numpy(eye_soa).to(dev)
        self.plasticDeviationB = torch.zeros(9 * T, dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        self._current = {"fp", "q"}

    # ---- the two state forms --------------------------------------------------------------------
    def _no_fp(self, what):
        raise RuntimeError(f"{what} needs Fp, which the compact viscoplastic state (Gd) does not hold: "
                           "use compactState=False")

    def state_form(self, fused: bool):
        if self.compact:
            if not fused:
                self._no_fp("the per-plugin viscoplastic step")
            return
        want = "q" if fused else "fp"
        if want not in self._current:
            self._convert(want)
        self._current = {want}

    def _convert(self, want):
        T = self.simulation.numTetra
        if want == "q":
            _lib.call("nfx_vp_q_from_fp", openCL.handle(), self.plasticDeformationTensorB.data_ptr(), T,
                      self.plasticDeviationB.data_ptr())
            self.simulation.plastic_state_changed()
        else:
            _lib.call("nfx_vp_fp_from_q", openCL.handle(), self.plasticDeviationB.data_ptr(), T,
                      self.plasticDeformationTensorB.data_ptr())
        self._current.add(want)

    def sync_forms(self):
        if self.compact:
            return
        for want in ("fp", "q"):
            if want not in self._current:
                self._convert(want)

    def forms_restored(self):
        self._current = {"gd"} if self.compact else {"fp", "q"}

    def state_buffers(self):
        if self.compact:
            return [self.plasticDeviationSymB, self.accumulatedStrainB]
        return [self.plasticDeformationTensorB, self.plasticDeviationB, self.accumulatedStrainB]

    def plastic_deformation(self):
        """The plastic deformation gradient Fp (9T SoA device tensor), up to date."""
        if self.compact:
            self._no_fp("plastic_deformation()")
        self.sync_forms()
        openCL.getCommandQueue().finish()
        return self.plasticDeformationTensorB

    # ---- per-plugin step ------------------------------------------------------------------------
    def build(self):
        self.scope = self.simulation.numTetra
        self.knl = self._enqueue_force

    def _enqueue_force(self):
        sim, h = self.simulation, openCL.handle()
        self.state_form(fused=False)
        vm = sim.vonMises.ptr if self.write_stress else None
        pr = sim.pressure.ptr if self.write_stress else None
        args = (self.params, sim.points.ptr, sim.tetras.ptr, self.dminvB.data_ptr(), sim.numPoints, sim.numTetra,
                self.plasticDeformationTensorB.data_ptr(), self.accumulatedStrainB.data_ptr(), vm, pr)
        if getattr(sim, "scatter", "atomic") == "atomic":
            _lib.call("nfx_viscoplastic_force", h, *args, sim.force.ptr, None)
            return
        off, slots, corners = sim.node_csr()
        _lib.call("nfx_viscoplastic_force", h, *args, None, corners.data_ptr())
        _lib.call("nfx_gather_corners", h, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), sim.numPoints,
                  sim.force.ptr)

    # ---- fused step ----------------------------------------------------------------------------
    def contribute(self, plan) -> bool:
        if plan.dminv:
            return False
        plan.model = _lib.NFX_VISCOPLASTIC
        plan.vp = self.params
        plan.dminv, plan.v0 = self.dminvB.data_ptr(), self.volumesB.data_ptr()
        if self.compact:
            if getattr(self.simulation, "scatter", "atomic") != "tiled":
                self._no_fp(f"the {self.simulation.scatter} scatter's viscoplastic step")
            plan.vp_state, plan.vp_sym = self.plasticDeviationSymB.data_ptr(), 1
            plan.epsbar = self.accumulatedStrainB.data_ptr() if self._reads_strain() else 0
        else:
            plan.vp_state = self.plasticDeviationB.data_ptr()
            plan.epsbar = self.accumulatedStrainB.data_ptr()
        if self.write_stress:
            plan.von_mises, plan.pressure = self.simulation.vonMises.ptr, self.simulation.pressure.ptr
        return True
