"""Seeded benchmark cells in the reference's JSON config format.

Unit mapping (the reference's system, ``Simulation.py:42-56``): E_sim = 0.01, rho_sim = 1, dt = 1,
L0 = RadiusSI / RadiusSIM.  A physical density of 1000 kg/m^3 with E ~ 1 kPa gives T0 ~ 1e-7 s, so a
0.2 s indentation would need ~2M steps.  The step counts of the benchmark cells are reached by mass
scaling (a quasi-static reading): ``RhoSI`` is chosen so that T0 = (simulated time) / (steps).

``small_cell``  plain, CPU-runnable: R 8 um, lattice 0.8 um, seed 1 (~4.2k nodes / ~25k tetra),
                Neo-Hookean E 1 kPa, nu 0.45; substrate at y = -7.5 um; tip R 2.5 um at 10 um/s to
                2 um over 20 000 steps, recorded every 100 steps.
``large_cell``  plain, large: R 10 um, lattice 0.2 um, seed 2 (~0.52M nodes / ~3.3M tetra),
                Mooney-Rivlin C10 150 Pa, C01 50 Pa, K 5 kPa (mu1 300, mu2 100, ratio 0.75,
                E 1168.8 Pa, nu 0.4610) with the Maxwell-type viscoplastic mapping
                (m = 1, n = 0, Y = 3 eta * 1/s, rate = T0 per step, yield offset 200 Pa);
                tip R 5 um, 3 um at 5 um/s then a 2 s hold; T0 = 1e-5 s; the wall touches the
                cell's bottom (the reference's placement, -(RadiusSIM + InitialDistance)).
The reference's vertical axis is y; "bottom z < -7.5 um" maps to y.  ``InitialDistance`` sets
both the tip start and the wall in the reference (``Simulation.py:107,163``); the small cell's wall
offset is given separately (``SubstrateDistance``, read by ``setInteractionSubstrate``).  The
large cell keeps the reference's wall: 0.5 um inside its bottom is 2.5 of its 0.2 um lattice
spacings, and the sudden compression of that cap makes the explicit step of this nearly
incompressible material diverge within ten steps (the FP64 oracle does the same).
"""

from __future__ import annotations

import numpy as np

from nofluidx3d_amd.mesh import generate_sphere_mesh


def _rho_for(T0, L0, E_si):
    p0 = E_si / 1e-2
    V0 = L0 / T0
    return p0 / V0**2


def small_cell(outdir="/tmp/nfx_small/", radius_sim=20.0, steps=20000, frames=200):
    R_si, E_si = 8e-6, 1000.0
    L0 = R_si / radius_sim
    move = 2e-6 / 10e-6  # 0.2 s
    T0 = move / steps
    return {
        "gpu": 0,
        "RhoSI": _rho_for(T0, L0, E_si),
        "SimulationTimeSI": move,
        "VTKFramesTotal": frames,
        "OutputDir": outdir,
        "InitialDistance": 0.0,
        "SubstrateDistance": -0.5e-6 / L0,
        "MoveTimeSI": move,
        "VelocitySI": 10e-6,
        "SphereRadiusSI": 2.5e-6,
        "CELL": {
            "RadiusSI": R_si,
            "RadiusSIM": radius_sim,
            "YoungsModulusSI": E_si,
            "PoissonRatio": 0.45,
            "model": "NEO_HOOKEAN",
            "InitVTK": outdir + "cell.vtk",
        },
        "MESH": {"radius_um": 8.0, "spacing_um": 0.8, "seed": 1},
    }


def large_cell(outdir="/tmp/nfx_large/", radius_sim=50.0, T0=1e-5):
    R_si = 10e-6
    mu1, mu2, K = 300.0, 100.0, 5000.0
    mu = mu1 + mu2
    E_si = 9 * K * mu / (3 * K + mu)
    nu = (3 * K - 2 * mu) / (2 * (3 * K + mu))
    L0 = R_si / radius_sim
    move = 3e-6 / 5e-6
    total = move + 2.0
    return {
        "gpu": 0,
        "RhoSI": _rho_for(T0, L0, E_si),
        "SimulationTimeSI": total,
        "VTKFramesTotal": 260,
        "OutputDir": outdir,
        "InitialDistance": 0.0,
        "MoveTimeSI": move,
        "VelocitySI": 5e-6,
        "SphereRadiusSI": 5e-6,
        "CELL": {
            "RadiusSI": R_si,
            "RadiusSIM": radius_sim,
            "YoungsModulusSI": E_si,
            "PoissonRatio": nu,
            "MooneyRivlinRatio": mu1 / mu,
            "model": "MOONEY_RIVLIN",
            "ViscositySI": 100.0,
            "YieldStressSI": 200.0,
            "InitVTK": outdir + "cell.vtk",
        },
        "MESH": {"radius_um": 10.0, "spacing_um": 0.2, "seed": 2},
    }


def make_mesh(cfg):
    m = cfg["MESH"]
    return generate_sphere_mesh(m["radius_um"], m["spacing_um"], m["seed"])


def write_mesh(cfg, points=None, tets=None):
    """Write the config's seeded mesh to ``CELL.InitVTK`` (legacy VTK, binary)."""
    import os

    from nofluidx3d_amd.util import vtk

    if points is None:
        points, tets = make_mesh(cfg)
    os.makedirs(os.path.dirname(cfg["CELL"]["InitVTK"]), exist_ok=True)
    vtk.from_tetra(points * 1e-6, tets).save(cfg["CELL"]["InitVTK"])
    return points, tets
