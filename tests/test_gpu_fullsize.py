"""GPU parity at the benchmark workloads' full sizes against the FP64 oracle (its C restatement on
all host cores).

- the large cell (R 10 um, 0.2 um lattice, seed 2: 523 305 nodes, 3 377 961 tetra; Mooney-Rivlin
  + viscoplastic, tiled scatter, hipGraph replay) over its first 500 steps, the tip pressing in
  from step 0: positions within 1e-4 RadiusSIM and within 1 % of the largest displacement, the
  tip-force history within 5e-4 of its maximum (measured: 1.4e-6 of RadiusSIM, 4.3e-4 of the
  displacement, 4.0e-4 of max|F|); a second run is bitwise identical.  The force is k = 0.1
  times the penetrations of the few nodes inside the tip, a difference of coordinates near 75, so
  the FP32 rounding of those positions (~5e-6, 1e-7 of RadiusSIM) alone is a few 1e-4 of this
  early-contact force (max 0.009).
- the 512-cell batch (seeds 1000+i, 17 370 772 tetra) over its whole 10 000-step protocol, three
  cells (the first, a middle and the last), and the same cells on the plain single-cell path
  (another FP32 formulation: (x, v, f_old) state, the shape state S, its own force reduction).
  Over 10 000 viscoplastic steps an FP32 trajectory drifts from the FP64 one by a few 1e-4, so
  the bars here are 5e-4 of RadiusSIM for positions and 1.5e-3 of max|F| for the curves, for
  both paths.  Measured (positions / curves): batch 0.6-2.4e-4 / 2.7-6.3e-4, plain path
  0.2-1.0e-4 / 0.8-9.3e-4.  The batch's symmetric state G = S S^T accounts for its larger
  position drift (cell 511: 2.4e-4, against 3.8e-5 for the batch run on S, which tracks the
  plain path to 2.5e-6).
"""

import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd import openCL

    openCL.initializeOpenCLObjects(0)
    return openCL.handle()


def large_cell_sim():
    from nofluidx3d_amd import configs
    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    cfg = configs.large_cell(outdir="/tmp/nfx_test_large/")
    m = cfg["MESH"]
    path = f"/tmp/nfx_mesh_{m['radius_um']}_{m['spacing_um']}_{m['seed']}.npz"  # shared with bench.py
    if os.path.exists(path):
        d = np.load(path)
        P, T = d["P"], d["T"]
    else:
        P, T = configs.make_mesh(cfg)
        tmp = f"{path}.{os.getpid()}.tmp.npz"
        np.savez(tmp, P=P, T=T)
        os.replace(tmp, path)
    sim = Simulation(cfg, scatter="tiled", mesh=vtk.from_tetra(P, T), verbose=False, tile_k=512)
    sim.setInteractionSubstrate()
    sim.setInteractionSphere()
    sim.setInteractionFiniteStrainViscoplastic()
    sim.setInteractionVelocityVerlet()
    return sim


def test_large_cell_first_steps_match_oracle_and_repeat_bitwise(ctx):
    import torch

    from test_gpu_plain import host, oracle_for

    n = 500
    sim = large_cell_sim()
    assert (sim.numPoints, sim.numTetra) == (523305, 3377961)
    x0 = sim.positions().copy()
    hist = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    sim.advance(n, tip_hist=hist)
    gx, gf = sim.positions(), host(hist)
    ref = oracle_for(sim, nthreads=O.max_threads())
    of = ref.run(n)
    ox = ref.positions()
    R = sim.parameters["CELL"]["RadiusSIM"]
    assert np.all(np.isfinite(ox)) and np.all(np.isfinite(gx))
    pe = np.max(np.abs(gx - ox)) / R
    moved = np.max(np.abs(ox - x0))
    fe = np.max(np.abs(gf - of)) / np.max(np.abs(of))
    print(f"large cell {n} steps: pos_err {pe:.2e} of R, {np.max(np.abs(gx - ox)) / moved:.2e} of the largest "
          f"displacement {moved:.3e}; force_err {fe:.2e} of max|F| {np.max(np.abs(of)):.3e}")
    assert pe < 1e-4 and moved > 0
    assert np.max(np.abs(gx - ox)) < 1e-2 * moved
    assert np.max(np.abs(of)) > 0 and fe < 5e-4
    again = large_cell_sim()
    hist2 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    again.advance(n, tip_hist=hist2)
    assert np.array_equal(again.positions(), gx) and np.array_equal(host(hist2), gf)


def test_full_batch_cells_match_oracle(ctx):
    from nofluidx3d_amd.batch import BatchSimulation, make_meshes, sample_batch, sim_points

    steps, ncurve, n = 10000, 256, 10000
    specs, _ = sample_batch(512, seed0=1000, steps=steps)
    meshes = make_meshes(specs)
    b = BatchSimulation(specs, meshes, steps=steps, ncurve=ncurve, tile_k=256)
    assert b.T_real == 17370772
    b.advance(n)
    curves = b.curves_sim()
    s_all = np.arange(n)
    k = s_all * ncurve // steps
    sampled = s_all[np.concatenate([[True], k[1:] != k[:-1]])]  # the steps whose force a curve point holds
    errs = {}
    for i in (0, 255, 511):
        s, c = b.specs[i], b.cells_host[i]
        prm = O.RunParams()
        prm.model = O.VISCOPLASTIC
        prm.vp = O.VPParams(0.01 / (2 * (1 + s.nu)), 0.0, 0.01 / (3 * (1 - 2 * s.nu)), 1.0,
                            3.0 * s.viscosity_si / b.units[i]["p0"], b.units[i]["T0"], 1.0, 0.0, float(c["vp_offset"]))
        prm.substrate, prm.substrate_k, prm.wall = 1, 0.1, float(c["wall"])
        prm.indenter, prm.ind_k, prm.ind_radius = 1, 0.1, s.tip_radius_um * 1e-6 / b.units[i]["L0"]
        prm.ind_start, prm.ind_speed, prm.ind_move_steps = float(c["tip_start"]), float(c["tip_speed"]), float(steps)
        prm.nthreads = O.max_threads()
        P, T = meshes[i]
        st = O.OracleState(sim_points(P, b.radius_sim[i]), T, prm)
        hist = st.run(n)
        R = b.radius_sim[i]
        ref, ox = hist[sampled], st.positions()
        fmax = np.max(np.abs(ref))
        px, pf = plain_path_run(b, meshes, i, n)
        e = {"batch": (np.max(np.abs(b.positions(i) - ox)) / R, np.max(np.abs(curves[i][:len(sampled)] - ref)) / fmax),
             "plain": (np.max(np.abs(px - ox)) / R, np.max(np.abs(pf[sampled] - ref)) / fmax)}
        errs[i] = (e, fmax)
        print(f"cell {i} (RadiusSIM {R:.1f}, max|F| {fmax:.3e}): batch pos {e['batch'][0]:.2e} curve {e['batch'][1]:.2e}; "
              f"plain path pos {e['plain'][0]:.2e} curve {e['plain'][1]:.2e}")
    for i, (e, fmax) in errs.items():
        assert fmax > 0, i
        for path, (pos, curve) in e.items():
            assert pos < 5e-4 and curve < 1.5e-3, (i, path)


def plain_path_run(b, meshes, i, n):
    """cell i of batch b on the plain single-cell Simulation (fused tiled step): positions after n
    steps and the per-step FP64 tip force"""
    import torch

    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    P, T = meshes[i]
    sim = Simulation(b.single_cell_config(i), mesh=vtk.from_tetra(P, T), verbose=False, tile_k=b.tile_k)
    sim.setInteractionSubstrate()
    sim.setInteractionSphere()
    sim.setInteractionFiniteStrainViscoplastic()
    sim.setInteractionVelocityVerlet()
    hist = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    sim.advance(n, tip_hist=hist)
    return sim.positions(), hist.cpu().numpy()
