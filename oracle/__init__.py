"""TEST INFRASTRUCTURE ONLY -- the FP64 CPU oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use this
package, as the checker (or the timed CPU stand-in for the reference's CPU path); the product
(`nofluidx3d_amd`) never imports it.

Contents
- ctypes bindings of ``nfx_oracle.c`` (FP64 restatement of the reference's OpenCL kernels);
- numpy restatements of the host-side arithmetic the kernels consume:
  ``edge_vectors_volumes``  -- reference ``nofluidx3d/TetraCell.py:26-41``
  ``units``                 -- reference ``nofluidx3d/Simulation.py:42-56``
  ``indenter_position``     -- reference ``nofluidx3d/Simulation.py:117-132, 173-183``
  ``force_si_sphere``       -- reference ``nofluidx3d/Simulation.py:198-212``
- ``mlp`` (submodule): FP64 numpy MLP forward/backward/Adam, the checker of the surrogate kernels.

Pins (tests/test_oracle.py): closed-form stresses, energy consistency by central differences,
objectivity, momentum balance, the exact two-step integrator recurrence, and brute-force sums.
No function here is "parity unpinned".
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnfx_oracle.so")

SVKK, SVK, MR, VISCOPLASTIC = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_uint32)


class VPParams(ctypes.Structure):
    _fields_ = [
        ("mu1", ctypes.c_double),
        ("mu2", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("flow_exponent", ctypes.c_double),
        ("yield_stress", ctypes.c_double),
        ("flow_rate", ctypes.c_double),
        ("hardening_threshold", ctypes.c_double),
        ("hardening_exponent", ctypes.c_double),
        ("yield_offset", ctypes.c_double),
    ]


class LegacyParams(ctypes.Structure):
    """parameters of the legacy stress laws (orc_legacy_force): law 0 uses E, nu; law 1 shear,
    bulk; law 2 shear, bulk, shear2; law 3 shear, phi"""

    _fields_ = [
        ("E", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("shear", ctypes.c_double),
        ("bulk", ctypes.c_double),
        ("shear2", ctypes.c_double),
        ("phi", ctypes.c_double),
    ]


LINEAR_ELASTIC, LINEAR_DEVIATORIC, NEO_HOOKEAN_2, PORO_NEO_HOOKEAN = 4, 5, 6, 7  # run models (law + 4)


class LVEParams(ctypes.Structure):
    """linear viscoelastic (Prony) law: E, nu, up to 8 terms of weights and relaxation times [steps]"""

    _fields_ = [("E", ctypes.c_double), ("nu", ctypes.c_double), ("nterms", ctypes.c_int32),
                ("weights", ctypes.c_double * 8), ("taus", ctypes.c_double * 8)]


def lve_params(E, nu, weights, taus):
    p = LVEParams()
    p.E, p.nu, p.nterms = E, nu, len(weights)
    for j, (w, t) in enumerate(zip(weights, taus)):
        p.weights[j], p.taus[j] = w, t
    return p


class RunParams(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int),
        ("E", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("ratio", ctypes.c_double),
        ("vp", VPParams),
        ("substrate", ctypes.c_int),
        ("substrate_k", ctypes.c_double),
        ("wall", ctypes.c_double),
        ("indenter", ctypes.c_int),
        ("ind_k", ctypes.c_double),
        ("ind_radius", ctypes.c_double),
        ("ind_start", ctypes.c_double),
        ("ind_speed", ctypes.c_double),
        ("ind_move_steps", ctypes.c_double),
        ("fix_top_bottom", ctypes.c_int),
        ("nthreads", ctypes.c_int),
        ("legacy", LegacyParams),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle.build import build

            build()
        L = ctypes.CDLL(LIB_PATH)
        i64 = ctypes.c_int64
        L.orc_tetra_force.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      i64, i64, _dp, _up, _dp, _dp, _dp, ctypes.c_int]
        L.orc_viscoplastic.argtypes = [ctypes.POINTER(VPParams), i64, i64, _dp, _up, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _dp, ctypes.c_int]
        L.orc_legacy_force.argtypes = [ctypes.c_int, ctypes.POINTER(LegacyParams), i64, i64, _dp, _up, _dp, _dp,
                                       _dp, _dp, ctypes.c_int]
        L.orc_sphere.argtypes = [i64, ctypes.c_double, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_linear_viscoelastic.argtypes = [ctypes.POINTER(LVEParams), i64, i64, _dp, _up, _dp, _dp, _dp, _dp, _dp,
                                              _dp]
        f64, _ip = ctypes.c_double, ctypes.POINTER(ctypes.c_int32)
        L.orc_round_tip.argtypes = [i64, _dp, _dp, f64, f64]
        L.orc_adhesive_plane.argtypes = [i64, _dp, _dp, f64, f64, f64, _ip]
        L.orc_point_zero_force.argtypes = [i64, _dp, _dp, f64]
        L.orc_center_of_mass.argtypes = [i64, _dp, _dp]
        L.orc_com_force.argtypes = [i64, _dp, f64, _dp]
        L.orc_over_damped.argtypes = [i64, ctypes.c_int, _dp, _dp]
        L.orc_velocity_verlet2.argtypes = [i64, _dp, _dp, _dp, _dp, _dp, f64]
        L.orc_volume_quadrature.argtypes = [ctypes.c_int, i64, i64, _dp, _dp, _up, f64, f64, f64, f64, f64, f64]
        L.orc_adhesive_plane_surface.argtypes = [i64, i64, _dp, _dp, _up, f64, f64, f64]
        L.orc_hard_adhesive_plane_surface.argtypes = [i64, i64, _dp, _dp, _up, f64, f64, f64]
        L.orc_substrate.argtypes = [i64, ctypes.c_double, _dp, _dp, ctypes.c_double]
        L.orc_plane.argtypes = [i64, ctypes.c_double, _dp, _dp, ctypes.c_double]
        L.orc_velocity_verlet.argtypes = [i64, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_tip_force.argtypes = [i64, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, _dp, _dp]
        L.orc_run.argtypes = [ctypes.POINTER(RunParams), i64, i64, _dp, _up, _dp, _dp, _dp, _dp, _dp,
                              _dp, _dp, i64, i64, _dp]
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _u(a):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_up)


def _dn(a):
    return None if a is None else _d(a)


# ---------------------------------------------------------------------------------------------
# numpy restatements of the host side
# ---------------------------------------------------------------------------------------------


def edge_vectors_volumes(points: np.ndarray, tets: np.ndarray):
    """TetraCell.py:26-41: edge vectors (T,9) = [p1-p0 | p2-p0 | p3-p0] row-major per edge, and
    signed volumes det/6."""
    p = points[tets]
    R1, R2, R3 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]
    ev = np.stack([R1, R2, R3], 1).reshape(len(tets), 9)
    vol = np.einsum("ij,ij->i", R1, np.cross(R2, R3)) / 6
    return ev, vol


def soa(a: np.ndarray, dtype=np.float64) -> np.ndarray:
    """(n, k) array to the reference's SoA buffer: element c of item i at c*n + i
    (``reshape(k*n, order="F")``, Simulation.py:82-87, Tetra.py:53-58)."""
    return np.ascontiguousarray(np.asarray(a).reshape(-1, order="F").astype(dtype))


def units(radius_si, radius_sim, youngs_si, rho_si):
    """Simulation.py:42-56 -> dict(p0, L0, V0, T0, rho0)."""
    rho0 = rho_si / 1.0
    L0 = radius_si / radius_sim
    p0 = youngs_si / 1e-2
    V0 = np.sqrt(p0 / rho0)
    return {"rho0": rho0, "L0": L0, "p0": p0, "V0": V0, "T0": L0 / V0}


def indenter_position(time, start, speed_per_step, move_steps):
    """Simulation.py:173-183 (sphere) / :117-132 (plane)."""
    if time <= move_steps:
        return start - speed_per_step * time
    return start - speed_per_step * move_steps


def force_si_sphere(points_n3: np.ndarray, sphere_y: float, radius: float, k: float, p0: float, L0: float):
    """Simulation.py:198-212 restated (numpy, FP64)."""
    c = np.array([0.0, sphere_y, 0.0])
    length = np.linalg.norm(c - points_n3, axis=-1)
    with np.errstate(invalid="ignore", divide="ignore"):
        normals = (points_n3 - c) / length[:, None]
    dis = np.maximum(radius - length, 0.0)
    force = (k * dis)[:, None] * normals * (p0 * L0**2)
    return -np.nansum(force[:, 1])


# ---------------------------------------------------------------------------------------------
# kernel wrappers (in-place on SoA float64 buffers)
# ---------------------------------------------------------------------------------------------


def tetra_force(model, E, nu, ratio, points, tetras, ev, vol, force, nthreads=1):
    N = points.size // 3
    T = tetras.size // 4
    lib().orc_tetra_force(model, E, nu, ratio, N, T, _d(points), _u(tetras), _d(ev), _d(vol), _d(force), nthreads)


def viscoplastic(vp: VPParams, points, tetras, ev, vol, force, acc, Fp, vonMises=None, pressure=None, nthreads=1):
    N = points.size // 3
    T = tetras.size // 4
    lib().orc_viscoplastic(ctypes.byref(vp), N, T, _d(points), _u(tetras), _d(ev), _d(vol), _d(force), _d(acc),
                           _d(Fp), _dn(vonMises), _dn(pressure), nthreads)


def legacy_force(law, P: LegacyParams, points, tetras, ev, force, vonMises=None, pressure=None, nthreads=1):
    """law 0..3 (linear elastic, linear deviatoric, second-order Neo-Hookean, poro-Neo-Hookean)"""
    N = points.size // 3
    T = tetras.size // 4
    lib().orc_legacy_force(law, ctypes.byref(P), N, T, _d(points), _u(tetras), _d(ev), _d(force), _dn(vonMises),
                           _dn(pressure), nthreads)


def _i32(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def round_tip(force, points, radius, pos):
    lib().orc_round_tip(points.size // 3, _d(force), _d(points), radius, pos)


def adhesive_plane(force, points, wall, adhesion, distance, on_surface):
    lib().orc_adhesive_plane(points.size // 3, _d(force), _d(points), wall, adhesion, distance, _i32(on_surface))


def point_zero_force(force, points, k):
    lib().orc_point_zero_force(points.size // 3, _d(force), _d(points), k)


def center_of_mass(points):
    com = np.zeros(2)
    lib().orc_center_of_mass(points.size // 3, _d(points), _d(com))
    return com


def com_force(force, k, com):
    lib().orc_com_force(force.size // 3, _d(force), k, _d(np.ascontiguousarray(com, np.float64)))


def over_damped(points, force, point3fix):
    lib().orc_over_damped(points.size // 3, int(point3fix), _d(points), _d(force))


def velocity_verlet2(points, vel, force, fold, masses, damping):
    lib().orc_velocity_verlet2(points.size // 3, _d(points), _d(vel), _d(force), _d(fold), _d(masses), damping)


def volume_quadrature(tip_shape, force, points, tetras, radius, half_angle, tip_pos, k):
    lib().orc_volume_quadrature(int(tip_shape), points.size // 3, tetras.size // 4, _d(force), _d(points),
                                _u(tetras), radius, half_angle, tip_pos[0], tip_pos[1], tip_pos[2], k)


def adhesive_plane_surface(force, points, triangles, wall, adhesion, distance):
    lib().orc_adhesive_plane_surface(points.size // 3, triangles.size // 3, _d(force), _d(points), _u(triangles),
                                     wall, adhesion, distance)


def hard_adhesive_plane_surface(force, points, triangles, wall, adhesion, force_const):
    lib().orc_hard_adhesive_plane_surface(points.size // 3, triangles.size // 3, _d(force), _d(points), _u(triangles),
                                          wall, adhesion, force_const)


def linear_viscoelastic(P: LVEParams, points, tetras, ev, force, devold, transient, vonMises=None, pressure=None):
    lib().orc_linear_viscoelastic(ctypes.byref(P), points.size // 3, tetras.size // 4, _d(points), _u(tetras), _d(ev),
                                  _d(force), _d(devold), _d(transient), _dn(vonMises), _dn(pressure))


def sphere(k, force, points, radius, pos):
    lib().orc_sphere(points.size // 3, k, _d(force), _d(points), radius, pos)


def substrate(k, force, points, wall):
    lib().orc_substrate(points.size // 3, k, _d(force), _d(points), wall)


def plane(k, force, points, wall):
    lib().orc_plane(points.size // 3, k, _d(force), _d(points), wall)


def velocity_verlet(points, vel, force, fold, fix_top_bottom=False):
    lib().orc_velocity_verlet(points.size // 3, int(fix_top_bottom), _d(points), _d(vel), _d(force), _d(fold))


def tip_force(points, indenter, k, radius, pos):
    f = ctypes.c_double()
    m = ctypes.c_double()
    lib().orc_tip_force(points.size // 3, _d(points), indenter, k, radius, pos, ctypes.byref(f), ctypes.byref(m))
    return f.value, m.value


def max_threads():
    return lib().orc_max_threads()


class OracleState:
    """FP64 state of one cell in the reference's SoA buffers, stepped by ``orc_run``."""

    def __init__(self, points_n3, tets_t4, params: RunParams):
        self.N = len(points_n3)
        self.T = len(tets_t4)
        ev, vol = edge_vectors_volumes(np.asarray(points_n3, np.float64), np.asarray(tets_t4))
        self.points = soa(points_n3)
        self.tetras = soa(tets_t4, np.uint32)
        self.ev = soa(ev)
        self.vol = np.ascontiguousarray(vol, dtype=np.float64)
        self.vel = np.zeros(3 * self.N)
        self.force = np.zeros(3 * self.N)
        self.fold = np.zeros(3 * self.N)
        self.acc = np.zeros(self.T)
        self.Fp = soa(np.tile(np.eye(3).reshape(1, 9), (self.T, 1)))
        self.params = params
        self.time = 0

    def run(self, nsteps, record=True):
        hist = np.zeros(nsteps) if record else None
        lib().orc_run(ctypes.byref(self.params), self.N, self.T, _d(self.points), _u(self.tetras), _d(self.ev),
                      _d(self.vol), _d(self.vel), _d(self.force), _d(self.fold), _d(self.acc), _d(self.Fp),
                      self.time, nsteps, _dn(hist))
        self.time += nsteps
        return hist

    def positions(self):
        return self.points.reshape(3, self.N).T.copy()
