"""GPU tests of the batched stepper (cells of different meshes and parameters in one launch per pass).

- every cell of a batch matches its own single-cell Simulation run (same force kernels, cell-aligned
  tiles): positions within 1e-5 * RadiusSIM, curve within 1e-4 * max|F| (the batch's (u, w)
  velocity-Verlet form and symmetric plastic deviation Gd round differently from the single cell's
  (u, v, f_old) form and Q);
- one cell against the FP64 oracle: 1e-4 (the plain-path tolerance);
- bitwise: repeated runs, graph vs eager;
- tile size (an autotuned choice) changes only the rounding: 1e-4;
- cache policy and the unread plastic strain: bitwise.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 800
NCURVE = 32


@pytest.fixture(scope="module")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd import openCL

    openCL.initializeOpenCLObjects(0)
    return openCL.handle()


def specs_for(viscous):
    from nofluidx3d_amd.batch import CellSpec

    geo = [(3.0, 0.6), (3.5, 0.7), (2.5, 0.5), (4.0, 0.8), (3.2, 0.8)]
    E = [300.0, 1000.0, 3000.0, 150.0, 800.0]
    nu = [0.45, 0.3, 0.49, 0.4, 0.35]
    eta = [40.0, 200.0, 900.0, 20.0, 100.0]
    tip = [1.0, 2.0, 1.5, 3.0, 1.2]
    v = [5.0, 0.8, 20.0, 2.0, 10.0]
    out = []
    for i in range(5):
        s = CellSpec(geo[i][0], geo[i][1], 3000 + i, E[i], nu[i], eta[i] if viscous else None, tip[i], v[i])
        assert s.relaxation_per_step(STEPS) <= 0.1
        out.append(s)
    return out


def make_batch(viscous, tile_k=256, use_graph=True, **kw):
    from nofluidx3d_amd.batch import BatchSimulation, make_meshes

    specs = specs_for(viscous)
    meshes = make_meshes(specs, workers=1, cache=None)
    return BatchSimulation(specs, meshes, steps=STEPS, ncurve=NCURVE, tile_k=tile_k, use_graph=use_graph, **kw), meshes


def sample_steps():
    s = np.arange(STEPS)
    k = s * NCURVE // STEPS
    first = np.concatenate([[True], k[1:] != k[:-1]])
    return s[first]


def single_cell(b, meshes, i):
    import torch

    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    cfg = b.single_cell_config(i)
    P, T = meshes[i]
    sim = Simulation(cfg, mesh=vtk.from_tetra(P, T), verbose=False, tile_k=b.tile_k)
    sim.setInteractionSubstrate()
    sim.setInteractionSphere()
    if b.specs[i].viscosity_si is not None:
        sim.setInteractionFiniteStrainViscoplastic()
    else:
        sim.setInteractionTetra()
    sim.setInteractionVelocityVerlet()
    hist = torch.zeros(STEPS, dtype=torch.float64, device="cuda:0")
    sim.advance(STEPS, tip_hist=hist)
    return sim, sim.positions(), hist.cpu().numpy()


@pytest.mark.parametrize("viscous", [False, True])
def test_batch_cells_match_single_cell_runs(ctx, viscous):
    b, meshes = make_batch(viscous)
    b.advance(STEPS)
    curves = b.curves_sim()
    ks = sample_steps()
    for i in range(b.ncells):
        sim, x, hist = single_cell(b, meshes, i)
        R = b.radius_sim[i]
        assert np.max(np.abs(b.positions(i) - x)) < 1e-5 * R, i
        ref = hist[ks]
        assert np.max(np.abs(ref)) > 0, i
        assert np.max(np.abs(curves[i] - ref)) < 1e-4 * np.max(np.abs(ref)), i


def test_batch_cell_matches_oracle(ctx):
    b, meshes = make_batch(True)
    b.advance(STEPS)
    i = 1
    s, c = b.specs[i], b.cells_host[i]
    prm = O.RunParams()
    prm.model = O.VISCOPLASTIC
    prm.vp = O.VPParams(float(c["mu1"]), float(c["mu2"]), float(c["kappa"]), 1.0, float(c["vp_yield"]),
                        float(c["vp_rate"]), 1.0, 0.0, float(c["vp_offset"]))
    prm.vp.mu1 = 0.01 / (2 * (1 + s.nu))
    prm.vp.kappa = 0.01 / (3 * (1 - 2 * s.nu))
    prm.vp.flow_rate = b.units[i]["T0"]
    prm.vp.yield_stress = 3.0 * s.viscosity_si / b.units[i]["p0"]
    prm.substrate, prm.substrate_k, prm.wall = 1, 0.1, float(c["wall"])
    prm.indenter, prm.ind_k, prm.ind_radius = 1, 0.1, s.tip_radius_um * 1e-6 / b.units[i]["L0"]
    prm.ind_start, prm.ind_speed, prm.ind_move_steps = float(c["tip_start"]), float(c["tip_speed"]), float(STEPS)
    prm.nthreads = 1
    from nofluidx3d_amd.batch import sim_points

    P, T = meshes[i]
    st = O.OracleState(sim_points(P, b.radius_sim[i]), T, prm)
    hist = st.run(STEPS)
    ref = hist[sample_steps()]
    R = b.radius_sim[i]
    assert np.max(np.abs(b.positions(i) - st.positions())) < 1e-4 * R
    assert np.max(np.abs(b.curves_sim()[i] - ref)) < 1e-4 * np.max(np.abs(ref))


def test_batch_bitwise_reproducible_graph_and_eager(ctx):
    a, _ = make_batch(True, use_graph=True)
    a.advance(300)
    b, _ = make_batch(True, use_graph=True)
    b.advance(300)
    c, _ = make_batch(True, use_graph=False)
    c.advance(300)
    xa, xb, xc = (z.x.cpu().numpy() for z in (a, b, c))
    assert np.array_equal(xa, xb) and np.array_equal(xa, xc)
    assert np.array_equal(a.curves_sim(), b.curves_sim()) and np.array_equal(a.curves_sim(), c.curves_sim())


def test_tile_size_changes_only_rounding(ctx):
    a, _ = make_batch(False, tile_k=256)
    a.advance(STEPS)
    b, _ = make_batch(False, tile_k=512)
    b.advance(STEPS)
    ca, cb = a.curves_sim(), b.curves_sim()
    assert np.max(np.abs(ca - cb)) < 1e-4 * np.max(np.abs(ca))
    for i in range(a.ncells):
        assert np.max(np.abs(a.positions(i) - b.positions(i))) < 1e-4 * a.radius_sim[i]


def test_cache_policy_and_plastic_strain_change_nothing(ctx):
    """The nontemporal cache policy of the force pass and keeping the (unread) plastic strain of the
    linear Maxwell law leave every bit of the run unchanged."""
    a, _ = make_batch(True, stream_nt=0, plastic_strain=True)
    a.advance(STEPS)
    b, _ = make_batch(True, stream_nt=1, plastic_strain=False)
    b.advance(STEPS)
    assert a.eps is not None and b.eps is None
    assert float(a.eps.abs().max()) > 0
    assert np.array_equal(a.curves_sim(), b.curves_sim())
    for i in range(a.ncells):
        assert np.array_equal(a.positions(i), b.positions(i))



def test_symmetric_deviation_matches_full_deviation(ctx):
    """The batch's symmetric plastic deviation Gd = Fp^-1 Fp^-T - I against Q = Fp^-1 - I: the same
    dynamics (the stress and flow depend on Gd alone), rounding apart."""
    a, _ = make_batch(True, sym_state=True)
    a.advance(STEPS)
    b, _ = make_batch(True, sym_state=False)
    b.advance(STEPS)
    assert a.state.numel() == 6 * a.T and b.state.numel() == 9 * b.T
    ca, cb = a.curves_sim(), b.curves_sim()
    assert np.max(np.abs(cb)) > 0
    assert np.max(np.abs(ca - cb)) < 1e-4 * np.max(np.abs(cb))
    for i in range(a.ncells):
        assert np.max(np.abs(a.positions(i) - b.positions(i))) < 1e-5 * a.radius_sim[i]


def test_shape_from_reference_positions_matches_streamed_shape(ctx):
    """The force pass forming Dm^-1 from the FP32 reference positions against the one streaming the
    FP64-precomputed Dm^-1: the same reference shape, FP32 rounding apart."""
    a, _ = make_batch(True, shape_from_ref=True)
    a.advance(STEPS)
    b, _ = make_batch(True, shape_from_ref=False)
    b.advance(STEPS)
    assert a.plan.shape_from_ref == 1 and b.plan.shape_from_ref == 0
    ca, cb = a.curves_sim(), b.curves_sim()
    assert np.max(np.abs(cb)) > 0
    assert np.max(np.abs(ca - cb)) < 1e-5 * np.max(np.abs(cb))
    for i in range(a.ncells):
        assert np.max(np.abs(a.positions(i) - b.positions(i))) < 1e-5 * a.radius_sim[i]


def test_narrow_tile_metadata_changes_no_bit(ctx):
    """16-bit tile-node ids and partial positions against the 32-bit arrays: the same run, bitwise."""
    a, _ = make_batch(True)
    assert a.t_narrow is not None
    a.advance(STEPS)
    b, _ = make_batch(True, narrow=False)
    assert b.t_narrow is None
    b.advance(STEPS)
    assert np.array_equal(a.curves_sim(), b.curves_sim())
    assert np.array_equal(a.x.cpu().numpy(), b.x.cpu().numpy())


def test_reset_reproduces_a_fresh_run(ctx):
    """reset() rebuilds the initial state on the device: a rerun equals the first run bitwise, and
    new physics equals a batch built with it."""
    import dataclasses

    a, meshes = make_batch(True)
    first_radius = a.radius_sim.copy()
    a.advance(STEPS)
    first = a.curves_sim().copy()
    xa = a.x.cpu().numpy().copy()
    a.reset(a.specs)
    a.advance(STEPS)
    assert np.array_equal(a.curves_sim(), first) and np.array_equal(a.x.cpu().numpy(), xa)
    # new E, viscosity, tip and nu (the relaxation per step stays put): nu moves RadiusSIM, so the
    # geometry is rebuilt on the device, exactly as construction builds it
    new = [dataclasses.replace(s, E_si=1.3 * s.E_si, viscosity_si=1.3 * s.viscosity_si, nu=0.97 * s.nu,
                               tip_radius_um=1.1 * s.tip_radius_um) for s in a.specs]
    from nofluidx3d_amd.batch import BatchSimulation

    b = BatchSimulation(new, meshes, steps=STEPS, ncurve=NCURVE, tile_k=256)
    a.reset(new)
    assert not np.array_equal(a.radius_sim, first_radius)
    assert np.array_equal(a.x.cpu().numpy(), b.x.cpu().numpy())
    assert np.array_equal(a.ref.cpu().numpy(), b.ref.cpu().numpy())
    assert np.array_equal(a.dminv.cpu().numpy(), b.dminv.cpu().numpy())
    assert np.array_equal(a.v0.cpu().numpy(), b.v0.cpu().numpy())
    assert np.array_equal(a.state.cpu().numpy(), b.state.cpu().numpy())
    a.advance(STEPS)
    b.advance(STEPS)
    assert np.array_equal(a.curves_sim(), b.curves_sim())


def test_cell_groups_equal_one_batch_bitwise(ctx):
    """The cells run as consecutive groups (each group all its steps before the next) give every
    cell bitwise the positions and curve it has in one batch; reset keeps that."""
    from nofluidx3d_amd.batch import BatchSimulation, CellGroups, cell_groups, make_meshes

    specs = specs_for(True)
    meshes = make_meshes(specs, workers=1, cache=None)
    sizes = [len(m[1]) for m in meshes]
    budget = sizes[0] + sizes[1]  # groups [0, 2), then by size
    assert len(cell_groups(meshes, budget)) >= 3
    one = BatchSimulation(specs, meshes, steps=STEPS, ncurve=NCURVE, tile_k=256, stream_nt=0)
    grp = CellGroups(specs, meshes, max_tetra=budget, steps=STEPS, ncurve=NCURVE, tile_k=256, stream_nt=0)
    assert grp.ranges == cell_groups(meshes, budget) and grp.T_real == one.T_real
    for b in (one, grp):
        b.advance(STEPS)
    assert np.array_equal(grp.curves_sim(), one.curves_sim())
    for i in range(len(specs)):
        assert np.array_equal(grp.positions(i), one.positions(i)), i
    assert np.array_equal(grp.radius_sim, one.radius_sim) and grp.cells_host.tobytes() == one.cells_host.tobytes()
    import dataclasses

    new = [dataclasses.replace(s, E_si=1.2 * s.E_si, viscosity_si=1.2 * s.viscosity_si) for s in specs]
    for b in (one, grp):
        b.reset(new)
        b.advance(STEPS)
    assert np.array_equal(grp.curves_sim(), one.curves_sim())

