"""Kernel registry with the API of reference ``nofluidx3d/KernelBuilder.py:25-62``, for plugins
written against the reference (the built-in plugins bind their libnfx entry points directly).

The reference concatenates process-global ``#define``s with shared headers and a ``.cl`` file and
JIT-compiles it per plugin.  Here every kernel is precompiled for gfx950 inside libnfx.so;
``KernelBuilder.build(file, kernelName)`` returns a :class:`Kernel` bound to the libnfx entry point
that implements ``kernelName``, carrying a snapshot of the current defines (mesh sizes, material
constants, model flags) as its launch parameters.  ``define``/``getDefines`` behave as in the
reference (values persist process-wide; booleans render lower-case).
"""

from __future__ import annotations

from pathlib import Path
from typing import Any

from nofluidx3d_amd import _lib, openCL


class Kernel:
    """A precompiled kernel; ``set_args`` takes the reference kernel's argument list."""

    def __init__(self, name: str, defines: set, values: dict):
        self.name = name
        self.defines = set(defines)
        self.values = dict(values)
        self.args = ()

    def set_args(self, *args):
        self.args = args

    def _v(self, key, default=None):
        v = self.values.get(key, default)
        if v is None:
            raise _lib.NfxError(f"kernel {self.name}: #define {key} is not set")
        return v

    def __call__(self):
        launch = _LAUNCHERS.get(self.name)
        if launch is None:
            raise _lib.NfxError(f"kernel {self.name!r} has no gfx950 implementation in libnfx")
        launch(self)


def _p(x):
    return x if isinstance(x, int) or x is None else x.data_ptr()


def _f(x):
    return float(getattr(x, "value", x))


_inverse_cache = {}


def _reference_inverse(edges, volumes, T):
    """(Dm^-1, V0) device pointers for calcTetraForce's reference-geometry arguments.  The
    reference's plugin passes its FP64 edge vectors (9T, SoA) and volumes (T) (``Tetra.py:49-58``):
    those are inverted once with the FP64 precompute and the result kept for the buffers' lifetime.
    FP32 arguments of the same sizes are taken as an inverse already (what ``Tetra`` itself holds).
    Anything else is refused: read as the wrong quantity, it would give wrong forces silently."""
    import torch

    if not (isinstance(edges, torch.Tensor) and isinstance(volumes, torch.Tensor)):
        raise _lib.NfxError("calcTetraForce: reference edge vectors and volumes must be device tensors")
    if edges.numel() != 9 * T or volumes.numel() != T:
        raise _lib.NfxError(f"calcTetraForce: expected 9T = {9 * T} edge-vector and T = {T} volume entries, got "
                            f"{edges.numel()} and {volumes.numel()}")
    if edges.dtype == torch.float32 and volumes.dtype == torch.float32:
        return edges.data_ptr(), volumes.data_ptr()
    if edges.dtype != torch.float64 or volumes.dtype != torch.float64:
        raise _lib.NfxError(f"calcTetraForce: edge vectors/volumes must both be float64 (the reference's buffers) or "
                            f"both float32 (an inverse), got {edges.dtype} and {volumes.dtype}")
    key = (edges.data_ptr(), volumes.data_ptr(), T)
    hit = _inverse_cache.get(key)
    if hit is None:
        dminv = torch.empty(9 * T, dtype=torch.float32, device=edges.device)
        v0 = torch.empty(T, dtype=torch.float32, device=edges.device)
        torch.cuda.synchronize(edges.device)
        _lib.call("nfx_precompute_f64", openCL.handle(), edges.data_ptr(), volumes.data_ptr(), T, dminv.data_ptr(),
                  v0.data_ptr())
        hit = _inverse_cache[key] = (dminv, v0, edges, volumes)  # the sources stay referenced with their key
    return hit[0].data_ptr(), hit[1].data_ptr()


def _launch_tetra(k: Kernel):
    # calcTetraForce(force, points, tetras, referenceEdgeVectors, referenceVolumes)
    force, points, tetras, edges, volumes = k.args
    if k.values.get("TETRA_SAINT_VENANT_KIRCHHOFF_KELLNBERGER"):
        model = _lib.NFX_SVKK
    elif k.values.get("TETRA_SAINT_VENANT_KIRCHHOFF"):
        model = _lib.NFX_SVK
    elif k.values.get("TETRA_MOONEY_RIVLIN"):
        model = _lib.NFX_MR
    else:
        raise _lib.NfxError("calcTetraForce: no material model flag set")
    mat = _lib.Elastic(model, float(k._v("def_tetraYoungsModulus")), float(k._v("def_tetraPoissonRatio")),
                       float(k._v("def_tetraMooneyRivlinRatio", 1)))
    N, T = int(k._v("def_pointCount")), int(k._v("def_tetraCount"))
    dminv, v0 = _reference_inverse(edges, volumes, T)
    _lib.call("nfx_tetra_force", openCL.handle(), mat, _p(points), _p(tetras), dminv, v0, N, T, _p(force))


def _launch_sphere(k: Kernel):
    force, points, radius, pos = k.args
    _lib.call("nfx_sphere_force", openCL.handle(), _p(force), _p(points), int(k._v("INSERT_NUM_POINTS")),
              float(k._v("def_FORCE_CONST")), _f(radius), _f(pos))


def _launch_substrate(k: Kernel):
    force, points, wall = k.args
    _lib.call("nfx_substrate_force", openCL.handle(), _p(force), _p(points), int(k._v("INSERT_NUM_POINTS")),
              float(k._v("def_FORCE_CONST")), _f(wall))


def _launch_plane(k: Kernel):
    force, points, wall = k.args
    _lib.call("nfx_plane_force", openCL.handle(), _p(force), _p(points), int(k._v("INSERT_NUM_POINTS")),
              float(k._v("def_FORCE_CONST")), _f(wall))


def _launch_vv(k: Kernel):
    points, vel, force, fold = k.args
    _lib.call("nfx_velocity_verlet", openCL.handle(), _p(points), _p(vel), _p(force), _p(fold),
              int(k._v("INSERT_NUM_POINTS")), 1 if "FIXED_TOP_BOTTOM" in k.defines else 0)


def _launch_vp(k: Kernel):
    # (force, points, tetras, Dm^-1, params, accumulatedStrain, Fp, vonMises, pressure)
    force, points, tetras, dminv, vp, acc, Fp, vm, pr = k.args
    N, T = int(k._v("INSERT_NUM_POINTS")), int(k._v("INSERT_NUM_TETRAS"))
    _lib.call("nfx_viscoplastic_force", openCL.handle(), vp, _p(points), _p(tetras), _p(dminv), N, T, _p(Fp), _p(acc),
              _p(vm), _p(pr), _p(force), None)


_LAUNCHERS = {
    "calcTetraForce": _launch_tetra,
    "Interaction_Sphere": _launch_sphere,
    "Interaction_Substrate": _launch_substrate,
    "Interaction_Plane": _launch_plane,
    "VelocityVerlet": _launch_vv,
    "Interaction_FiniteStrainViscoplastic": _launch_vp,
}


class KernelBuilder:
    defines: set[str] = set()
    valuesDefines: dict[str, Any] = {}

    @staticmethod
    def define(*args, **kwargs):
        KernelBuilder.defines.update(set(args))
        KernelBuilder.valuesDefines.update(kwargs)

    @staticmethod
    def getDefines() -> str:
        defines = ""
        for d in KernelBuilder.defines:
            defines += f"#define {d}\n"
        for name, value in KernelBuilder.valuesDefines.items():
            if isinstance(value, bool):
                value = str(value).lower()
            defines += f"#define {name} {value}\n"
        return defines

    @staticmethod
    def build(file: Path, kernelName: str) -> Kernel:
        if kernelName not in _LAUNCHERS:
            raise _lib.NfxError(f"{Path(file).name}:{kernelName} has no gfx950 implementation in libnfx")
        return Kernel(kernelName, KernelBuilder.defines, KernelBuilder.valuesDefines)
