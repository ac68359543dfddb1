"""GPU parity at the benchmark workloads' full sizes against the FP64 oracle (its C restatement on
all host cores).  Bars: positions within 1e-4 of RadiusSIM, tip-force curves within 1e-4 of their
maximum (the force is zero before contact, so a pointwise relative error is undefined).

- The large cell (R 10 um, 0.2 um lattice, seed 2: 523 305 nodes, 3 377 961 tetra; Mooney-Rivlin +
  Maxwell viscoplastic, substrate, spherical tip):
  - developed contact: the tip driven at 8x the protocol's speed (40 um/s) for 2000 steps, 0.8 um
    into the cell, tiled scatter with hipGraph replay; a second run is bitwise identical; the same
    window with the compact viscoplastic state;
  - the three scatter modes over the first 600 steps of that run: float atomics and the CSR gather
    against the oracle and against the tiled run;
  - the whole protocol (3 um at 5 um/s, then a 2 s hold: 260 000 steps) on the GPU alone, checked
    by what the physics fixes: finite positions, a force rising through the indentation and relaxing
    monotonically through the hold to a positive residual.
- The 512-cell batch (seeds 1000+i, 17 370 772 tetra) over its whole 10 000-step protocol: three
  cells (the first, a middle and the last) against the oracle, and the same cells on the plain
  single-cell path.
"""

import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FAST = 8.0  # tip speed factor of the developed-contact window
WINDOW, SHORT = 2000, 600


@pytest.fixture(scope="module")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd import openCL

    openCL.initializeOpenCLObjects(0)
    return openCL.handle()


def large_mesh(cfg):
    from nofluidx3d_amd import configs

    m = cfg["MESH"]
    path = f"/tmp/nfx_mesh_{m['radius_um']}_{m['spacing_um']}_{m['seed']}.npz"  # shared with bench.py
    if os.path.exists(path):
        d = np.load(path)
        return d["P"], d["T"]
    P, T = configs.make_mesh(cfg)
    tmp = f"{path}.{os.getpid()}.tmp.npz"
    np.savez(tmp, P=P, T=T)
    os.replace(tmp, path)
    return P, T


def large_cell_sim(scatter="tiled", speed=1.0, compact=False, node_fused=None):
    from nofluidx3d_amd import configs
    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    cfg = configs.large_cell(outdir="/tmp/nfx_test_large/")
    cfg["VelocitySI"] *= speed
    cfg["MoveTimeSI"] /= speed
    P, T = large_mesh(cfg)
    sim = Simulation(cfg, scatter=scatter, mesh=vtk.from_tetra(P, T), verbose=False, tile_k=512,
                     fuse_node_pass=node_fused)
    sim.setInteractionSubstrate()
    sim.setInteractionSphere()
    sim.setInteractionFiniteStrainViscoplastic(compactState=compact)
    sim.setInteractionVelocityVerlet()
    return sim


def run(sim, n):
    import torch

    from test_gpu_plain import host

    hist = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    sim.advance(n, tip_hist=hist)
    return sim.positions(), host(hist)


@pytest.fixture(scope="module")
def large_oracle(ctx):
    """The oracle's developed-contact run: positions after SHORT and WINDOW steps, force history."""
    from test_gpu_plain import oracle_for

    sim = large_cell_sim(speed=FAST)
    assert (sim.numPoints, sim.numTetra) == (523305, 3377961)
    ref = oracle_for(sim, nthreads=O.max_threads())
    f_short = ref.run(SHORT)
    x_short = ref.positions()
    f_rest = ref.run(WINDOW - SHORT)
    return {"x0": sim.cell.points.copy(), "x_short": x_short, "x": ref.positions(), "f": np.concatenate([f_short, f_rest]),
            "R": sim.parameters["CELL"]["RadiusSIM"], "acc": ref.acc.max()}


def errors(gx, gf, ox, of, R):
    return np.max(np.abs(gx - ox)) / R, np.max(np.abs(gf - of)) / np.max(np.abs(of))


def test_large_cell_developed_contact_matches_oracle(large_oracle):
    ref = large_oracle
    sim = large_cell_sim(speed=FAST)
    gx, gf = run(sim, WINDOW)
    pe, fe = errors(gx, gf, ref["x"], ref["f"], ref["R"])
    depth = sim.interactions[1].sphereFunc(0) - sim.interactions[1].sphereFunc(WINDOW)
    print(f"large cell {WINDOW} steps, tip {depth * sim.L0 * 1e6:.2f} um deep: pos_err {pe:.2e} of R, "
          f"force_err {fe:.2e} of max|F| {np.max(np.abs(ref['f'])):.3e}; largest displacement "
          f"{np.max(np.abs(ref['x'] - ref['x0'])):.3f}, epsbar max {ref['acc']:.2e}")
    assert np.all(np.isfinite(gx)) and ref["acc"] > 0  # the viscoplastic flow is active
    assert np.max(np.abs(ref["f"])) > 0.3  # developed contact (60x the force 500 steps into the protocol)
    assert pe < 1e-4 and fe < 1e-4
    again = large_cell_sim(speed=FAST)
    gx2, gf2 = run(again, WINDOW)
    assert np.array_equal(gx2, gx) and np.array_equal(gf2, gf)


def test_large_cell_compact_state_matches_oracle(large_oracle):
    """The compact viscoplastic state (Gd, no plastic strain) over the developed-contact window."""
    ref = large_oracle
    sim = large_cell_sim(speed=FAST, compact=True)
    gx, gf = run(sim, WINDOW)
    assert sim.plan().vp_sym == 1
    pe, fe = errors(gx, gf, ref["x"], ref["f"], ref["R"])
    print(f"compact state, {WINDOW} steps: pos_err {pe:.2e} of R, force_err {fe:.2e}")
    assert pe < 1e-4 and fe < 1e-4


def test_large_cell_node_fused_step_matches_oracle(large_oracle):
    """The node-fused step (one launch per step; not the size default here) over the developed-contact
    window at full size."""
    ref = large_oracle
    sim = large_cell_sim(speed=FAST, node_fused=True)
    gx, gf = run(sim, WINDOW)
    assert sim.plan().disp_b
    pe, fe = errors(gx, gf, ref["x"], ref["f"], ref["R"])
    print(f"node-fused step, {WINDOW} steps: pos_err {pe:.2e} of R, force_err {fe:.2e}")
    assert pe < 1e-4 and fe < 1e-4


@pytest.mark.parametrize("scatter", ["atomic", "csr"])
def test_large_cell_scatter_modes_match_oracle_and_tiled(large_oracle, scatter):
    ref = large_oracle
    tiled = large_cell_sim(speed=FAST)
    tx, tf = run(tiled, SHORT)
    sim = large_cell_sim(scatter=scatter, speed=FAST)
    gx, gf = run(sim, SHORT)
    assert sim.scatter == scatter
    pe, fe = errors(gx, gf, ref["x_short"], ref["f"][:SHORT], ref["R"])
    pt, ft = errors(gx, gf, tx, tf, ref["R"])
    print(f"{scatter}, {SHORT} steps: vs oracle pos {pe:.2e} force {fe:.2e}; vs tiled pos {pt:.2e} force {ft:.2e}")
    assert pe < 1e-4 and fe < 1e-4
    assert pt < 1e-4 and ft < 1e-4


def test_large_cell_full_protocol_properties(ctx):
    sim = large_cell_sim()
    n = sim.Nsteps
    move = int(round(sim.parameters["MoveTimeSI"] / sim.T0))
    assert (n, move) == (260000, 60000)
    gx, gf = run(sim, n)
    assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gf))
    x0 = sim.cell.points
    assert np.max(np.linalg.norm(gx - x0, axis=1)) < sim.parameters["CELL"]["RadiusSIM"]
    # indentation: the force rises through the descent
    rise = gf[[move // 4, move // 2, 3 * move // 4, move]]
    assert np.all(np.diff(rise) > 0) and rise[0] > 0, rise
    # hold: stress relaxation, monotone (sampled every 10 000 steps) down to a positive residual
    hold = gf[move::10000]
    assert np.all(np.diff(hold) <= 1e-3 * hold[0]), hold
    assert 0 < hold[-1] < 0.9 * hold[0], hold
    print(f"full protocol: F(indent 1/4..1) {rise}, hold {hold[0]:.4f} -> {hold[-1]:.4f}")


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
        prm.indenter, prm.ind_k, prm.ind_radius = 1, 0.1, float(c["tip_radius"])
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
            assert pos < 1e-4 and curve < 1e-4, (i, path)


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
