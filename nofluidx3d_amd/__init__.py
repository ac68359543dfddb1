"""MI355X-native implementation of noFluidX3D's explicit tetrahedral-FEM time step.

Public API mirrors the reference package ``nofluidx3d``: ``Simulation``, ``TetraCell``,
``interactions`` (Tetra, Sphere, Substrate, Plane, VelocityVerlet, FiniteStrainViscoplastic),
``openCL`` (device/queue singletons), ``KernelBuilder`` and ``util.Bridge``.  Compute runs in
hand-written HIP kernels for gfx950 in ``libnfx.so`` (declared in ``include/nfx.h``).
"""

__version__ = "0.1.0"


def __getattr__(name):
    if name == "Simulation":
        from nofluidx3d_amd.Simulation import Simulation

        return Simulation
    if name == "TetraCell":
        from nofluidx3d_amd.TetraCell import TetraCell

        return TetraCell
    raise AttributeError(name)
