"""Legacy stress laws behind the current plugin protocol (reference ``Interactions.py:101-247`` and
``352-408``: ``InteractionLinearElastic``, ``InteractionLinearElasticDeviatoric``,
``InteractionPoroNeoHookean``, ``InteractionSecondOrderNeoHookean``; parameters as
``SimulationLegacy.py:696-840`` sets them).

Each law is a Cauchy stress from the deformation gradient F = r Dm^-1 with node 0 as the edge base
(the legacy kernels used node 3; ``TetraCell.py:33-38`` uses node 0), applied through the kernels'
face-area nodal forces; the per-tetra ``vonMises`` and ``pressure`` are written when the plugin's
``write_stress`` is on (the reference writes them every step; here on the steps whose state a
frame shows).  Parameters are uniform over the mesh, as the legacy wrappers' per-tetra arrays are.
"""

from nofluidx3d_amd import _lib, openCL
from nofluidx3d_amd.interactions.Interaction import Interaction
from nofluidx3d_amd.interactions.Tetra import precompute_reference


class LegacyElastic(Interaction):
    model = None

    def __init__(self, referenceCell, E=0.0, nu=0.0, shear=0.0, bulk=0.0, shear2=0.0, volume_fraction=0.0):
        self.referenceCell = referenceCell
        self.params = _lib.Elastic(self.model, float(E), float(nu), 1.0, float(shear), float(bulk), float(shear2),
                                   float(volume_fraction))
        self.write_stress = False

    def createBuffers(self):
        self.dminvB, self.volumesB = precompute_reference(self.referenceCell, self.simulation)
        self.referenceCell = None

    def build(self):
        self.scope = self.simulation.numTetra

    def enqueue(self):
        sim = self.simulation
        h = openCL.handle()
        vm = sim.vonMises.ptr if self.write_stress else None
        pr = sim.pressure.ptr if self.write_stress else None
        if getattr(sim, "scatter", "atomic") in ("csr", "tiled"):
            off, slots, corners = sim.node_csr()
            _lib.call("nfx_legacy_force", h, self.params, sim.points.ptr, sim.tetras.ptr, self.dminvB.data_ptr(),
                      sim.numPoints, sim.numTetra, vm, pr, None, corners.data_ptr())
            _lib.call("nfx_gather_corners", h, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), sim.numPoints,
                      sim.force.ptr)
        else:
            _lib.call("nfx_legacy_force", h, self.params, sim.points.ptr, sim.tetras.ptr, self.dminvB.data_ptr(),
                      sim.numPoints, sim.numTetra, vm, pr, sim.force.ptr, None)

    def contribute(self, plan) -> bool:
        if plan.dminv:
            return False
        p = self.params
        plan.model = self.model
        plan.E, plan.nu, plan.ratio = p.E, p.nu, 1.0
        plan.shear, plan.bulk, plan.shear2, plan.volume_fraction = p.shear, p.bulk, p.shear2, p.volume_fraction
        plan.dminv = self.dminvB.data_ptr()
        plan.v0 = self.volumesB.data_ptr()
        if self.write_stress:
            plan.von_mises, plan.pressure = self.simulation.vonMises.ptr, self.simulation.pressure.ptr
        return True


class LinearElastic(LegacyElastic):
    """sigma = 2 mu eps + lambda tr(eps) I (InteractionLinearElastic.cl)"""

    model = _lib.NFX_LINEAR_ELASTIC

    def __init__(self, referenceCell, youngsModulus, poissonRatio):
        super().__init__(referenceCell, E=youngsModulus, nu=poissonRatio)


class LinearElasticDeviatoric(LegacyElastic):
    """sigma = 2 mu dev(eps) + K tr(eps) I (InteractionLinearElasticDeviatoric.cl)"""

    model = _lib.NFX_LINEAR_DEVIATORIC

    def __init__(self, referenceCell, shearModulus, bulkModulus):
        super().__init__(referenceCell, shear=shearModulus, bulk=bulkModulus)


class SecondOrderNeoHookean(LegacyElastic):
    """sigma = J^-5/3 (mu + mu2 (I1bar - 3)) dev(B) + K (J - 1) I (InteractionSecondOrderNeoHookean.cl)"""

    model = _lib.NFX_NEO_HOOKEAN_2

    def __init__(self, referenceCell, shearModulus, shearModulusSO, bulkModulus):
        super().__init__(referenceCell, shear=shearModulus, shear2=shearModulusSO, bulk=bulkModulus)


class PoroNeoHookean(LegacyElastic):
    """porous Neo-Hookean of Danielsson, Parks and Boyce (2002) (InteractionPoroNeoHookean.cl)"""

    model = _lib.NFX_PORO_NEO_HOOKEAN

    def __init__(self, referenceCell, shearModulus, volumeFraction):
        super().__init__(referenceCell, shear=shearModulus, volume_fraction=volumeFraction)


class LinearViscoelastic(Interaction):
    """Small-strain generalized Maxwell law (reference ``Interactions.py:518-600``, kernel
    ``InteractionLinearViscoelastic.cl``): sigma = 2 mu w_inf dev(eps) + kappa tr(eps) I + sum_j h_j,
    h_j relaxing with exp(-1/tau_j) per step (tau_j in steps), w_inf = 1 - sum_j w_j.  State: the
    previous deviatoric strain and the transient stresses, zero at the start (the reference's
    wrapper allocates them READ_ONLY and names an undefined buffer, Interactions.py:566-596; here
    they are ordinary state)."""

    def __init__(self, referenceCell, youngsModulus, poissonRatio, weights, relaxationTimes):
        if len(weights) != len(relaxationTimes):
            raise ValueError("weights and relaxationTimes must be of equal number")
        if len(weights) > 8:
            raise ValueError("at most 8 viscoelastic terms")
        self.referenceCell = referenceCell
        p = _lib.Viscoelastic()
        p.E, p.nu, p.nterms = float(youngsModulus), float(poissonRatio), len(weights)
        for j, (w, t) in enumerate(zip(weights, relaxationTimes)):
            p.weights[j], p.taus[j] = float(w), float(t)
        self.params = p
        self.write_stress = False

    def createBuffers(self):
        import torch

        self.dminvB, self.volumesB = precompute_reference(self.referenceCell, self.simulation)
        self.referenceCell = None
        T, dev = self.simulation.numTetra, self.dminvB.device
        self.oldDeviatoricStrainB = torch.zeros(9 * T, dtype=torch.float32, device=dev)
        self.transientStressB = torch.zeros(max(9 * self.params.nterms * T, 1), dtype=torch.float32, device=dev)

    def build(self):
        self.scope = self.simulation.numTetra

    def enqueue(self):
        sim = self.simulation
        h = openCL.handle()
        vm = sim.vonMises.ptr if self.write_stress else None
        pr = sim.pressure.ptr if self.write_stress else None
        args = (h, self.params, sim.points.ptr, sim.tetras.ptr, self.dminvB.data_ptr(), sim.numPoints, sim.numTetra,
                self.oldDeviatoricStrainB.data_ptr(), self.transientStressB.data_ptr(), vm, pr)
        if getattr(sim, "scatter", "atomic") in ("csr", "tiled"):
            off, slots, corners = sim.node_csr()
            _lib.call("nfx_linear_viscoelastic_force", *args, None, corners.data_ptr())
            _lib.call("nfx_gather_corners", h, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), sim.numPoints,
                      sim.force.ptr)
        else:
            _lib.call("nfx_linear_viscoelastic_force", *args, sim.force.ptr, None)
