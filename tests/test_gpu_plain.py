"""GPU parity of the plain path: each libnfx kernel through the C ABI against the FP64 oracle on
the same seeded inputs, then fused multi-step runs.

Tolerances (FP32 device state vs FP64 oracle):
- single force evaluation: max |f_gpu - f_oracle| <= 2e-5 * max |f_oracle|;
- multi-step runs: positions within 1e-4 * RadiusSIM and the indenter force curve within
  1e-4 * max |F| (the curve error is normalised by the curve maximum because the force is zero
  before contact, so a pointwise relative error is undefined);
- deterministic (CSR) mode: bitwise identical across repeated runs and across graph/eager.
"""

import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)


@pytest.fixture(scope="module")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd import openCL

    openCL.initializeOpenCLObjects(0)
    return openCL.handle()


def dev(a, dtype=None):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    t = torch.from_numpy(a if dtype is None else a.astype(dtype)).to("cuda:0")
    torch.cuda.synchronize()
    return t


def host(t, dtype=np.float64):
    import torch

    from nofluidx3d_amd import _lib, openCL

    _lib.call("nfx_stream_sync", openCL.handle())
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(dtype)


def call(name, *args):
    from nofluidx3d_amd import _lib

    _lib.call(name, *args)


def small_mesh(radius=3.0, spacing=0.6, seed=3, scale=5.0):
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, T = generate_sphere_mesh(radius, spacing, seed)
    return P * scale, T


def deformed(P, amp=0.08, noise=0.02):
    """smooth nonlinear deformation plus small noise (keeps every tetra positively oriented)"""
    F = np.eye(3) + amp * RNG.uniform(-1, 1, (3, 3))
    x = P @ F.T + amp * 0.05 * P**2 / np.abs(P).max()
    return x + noise * RNG.normal(size=P.shape)


def gpu_precompute(ctx, P, T):
    import torch

    ev, vol = O.edge_vectors_volumes(P, T)
    dminv = torch.empty(9 * len(T), dtype=torch.float32, device="cuda:0")
    v0 = torch.empty(len(T), dtype=torch.float32, device="cuda:0")
    evd, vold = dev(O.soa(ev)), dev(vol)
    call("nfx_precompute_f64", ctx, evd.data_ptr(), vold.data_ptr(), len(T), dminv.data_ptr(), v0.data_ptr())
    return dminv, v0, ev, vol


# ------------------------------------------------------------------------------------------------
# single kernels
# ------------------------------------------------------------------------------------------------


def test_precompute_inverse_fp64(ctx):
    P, T = small_mesh()
    dminv, v0, ev, vol = gpu_precompute(ctx, P, T)
    D = host(dminv).reshape(9, -1).T.reshape(-1, 3, 3)
    Dm = ev.reshape(-1, 3, 3).transpose(0, 2, 1)  # columns R1, R2, R3
    ref = np.linalg.inv(Dm)
    assert np.max(np.abs(D - ref) / np.abs(ref).max(axis=(1, 2))[:, None, None]) < 2e-7
    assert np.array_equal(host(v0, np.float32), vol.astype(np.float32))


@pytest.mark.parametrize("model,ratio", [(O.MR, 1.0), (O.MR, 0.75), (O.SVK, 1.0), (O.SVKK, 1.0)])
def test_tetra_force_parity(ctx, model, ratio):
    import torch

    from nofluidx3d_amd import _lib

    P, T = small_mesh()
    N, nT = len(P), len(T)
    dminv, v0, ev, vol = gpu_precompute(ctx, P, T)
    x = deformed(P)
    x32 = x.astype(np.float32).astype(np.float64)
    E, nu = 0.01, 0.45
    f_ref = np.zeros(3 * N)
    O.tetra_force(model, E, nu, ratio, O.soa(x32), O.soa(T, np.uint32), O.soa(ev), np.ascontiguousarray(vol), f_ref)
    xd, td = dev(O.soa(x32), np.float32), dev(O.soa(T, np.uint32))
    f = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    mat = _lib.Elastic(model, E, nu, ratio)
    call("nfx_tetra_force", ctx, mat, xd.data_ptr(), td.data_ptr(), dminv.data_ptr(), v0.data_ptr(), N, nT, f.data_ptr())
    assert np.all(np.isfinite(f_ref))
    err = np.max(np.abs(host(f) - f_ref)) / np.max(np.abs(f_ref))
    assert err < 2e-5, err
    # per-corner records + deterministic gather give the same node forces
    corners = torch.zeros(12 * nT, dtype=torch.float32, device="cuda:0")
    call("nfx_tetra_corner_forces", ctx, mat, xd.data_ptr(), td.data_ptr(), dminv.data_ptr(), v0.data_ptr(), N, nT,
         corners.data_ptr())
    c = host(corners).reshape(nT, 4, 3)
    assert np.max(np.abs(c.sum(1))) < 1e-6 * np.max(np.abs(c))  # f0 = -(f1+f2+f3)
    off, slots = csr(T, N)
    g = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    call("nfx_gather_corners", ctx, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), N, g.data_ptr())
    err = np.max(np.abs(host(g) - f_ref)) / np.max(np.abs(f_ref))
    assert err < 2e-5, err


def test_tetra_force_matches_golden_fixture(ctx):
    """tests/golden/homogeneous_states.npz (closed-form stresses, see tests/golden/make_golden.py)
    through nfx_precompute_f64 + nfx_tetra_force, FP32: within 2e-5 of the largest force."""
    import os

    import torch

    from nofluidx3d_amd import _lib

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "homogeneous_states.npz"))
    X = g["X"]
    T = np.array([[0, 1, 2, 3]])
    dminv, v0, _, _ = gpu_precompute(ctx, X, T)
    td = dev(O.soa(T, np.uint32))
    for i in range(len(g["names"])):
        x = g["x"][i]
        xd = dev(O.soa(x), np.float32)
        fd = torch.zeros(12, dtype=torch.float32, device="cuda:0")
        mat = _lib.Elastic(_lib.NFX_MR, float(g["E"]), float(g["nu"]), float(g["ratio"][i]))
        call("nfx_tetra_force", ctx, mat, xd.data_ptr(), td.data_ptr(), dminv.data_ptr(), v0.data_ptr(), 4, 1,
             fd.data_ptr())
        f = host(fd).reshape(3, -1).T
        ref = g["forces"][i]
        assert np.max(np.abs(f - ref)) < 2e-5 * np.abs(ref).max(), g["names"][i]


def csr(T, N):
    off = np.zeros(N + 1, np.uint32)
    slots = np.zeros(4 * len(T), np.uint32)
    tets = O.soa(T, np.uint32)
    call("nfx_build_node_csr_host", tets.ctypes.data_as(ctypes.c_void_p), N, len(T),
         off.ctypes.data_as(ctypes.c_void_p), slots.ctypes.data_as(ctypes.c_void_p))
    return dev(off), dev(slots)


def test_node_kernels_parity(ctx):
    import torch

    N = 5000
    x = RNG.normal(scale=4.0, size=(N, 3))
    x32 = x.astype(np.float32).astype(np.float64)
    f0 = RNG.normal(scale=1e-3, size=3 * N).astype(np.float32).astype(np.float64)
    xs = O.soa(x32)
    # sphere + substrate + plane
    for name, fn, args in (("nfx_sphere_force", O.sphere, (0.1, 3.0, 1.5)), ("nfx_substrate_force", O.substrate,
                                                                            (0.1, -2.0)),
                           ("nfx_plane_force", O.plane, (0.1, 2.5))):
        fr = f0.copy()
        fn(args[0], fr, xs, *args[1:])
        fd = dev(f0, np.float32)
        xsd = dev(xs, np.float32)  # referenced until the kernel has run
        call(name, ctx, fd.data_ptr(), xsd.data_ptr(), N, *args)
        assert np.max(np.abs(host(fd) - fr)) < 1e-6, name
    # velocity Verlet, with fixTopBottom
    v = RNG.normal(scale=1e-2, size=3 * N).astype(np.float32).astype(np.float64)
    fo = RNG.normal(scale=1e-3, size=3 * N).astype(np.float32).astype(np.float64)
    for fix in (0, 1):
        xr, vr, fr, for_ = xs.copy(), v.copy(), f0.copy(), fo.copy()
        O.velocity_verlet(xr, vr, fr, for_, fix_top_bottom=bool(fix))
        xd, vd, fd, fod = (dev(a, np.float32) for a in (xs, v, f0, fo))
        call("nfx_velocity_verlet", ctx, xd.data_ptr(), vd.data_ptr(), fd.data_ptr(), fod.data_ptr(), N, fix)
        assert np.max(np.abs(host(xd) - xr)) < 4e-6
        assert np.max(np.abs(host(vd) - vr)) < 1e-7
        assert np.array_equal(host(fod), for_) and np.all(host(fd) == 0)
    # FP64 tip-force reduction vs oracle sum on the same FP32 positions, from absolute positions and
    # from displacements u against reference positions X0 (x = X0 + u summed in FP64)
    out = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    X0 = O.soa((x + 30.0).astype(np.float32).astype(np.float64))
    u = (xs - X0).astype(np.float32).astype(np.float64)
    for ind, R, pos in ((1, 3.0, 2.0), (2, 0.0, 1.0)):
        xsd = dev(xs, np.float32)
        call("nfx_tip_force", ctx, xsd.data_ptr(), None, N, ind, 0.1, R, pos, out.data_ptr())
        f, m = host(out)
        fr, mr = O.tip_force(xs, ind, 0.1, R, pos)
        assert f == pytest.approx(fr, rel=1e-12) and m == pytest.approx(mr, rel=1e-12)
        ud, X0d = dev(u, np.float32), dev(X0, np.float32)
        call("nfx_tip_force", ctx, ud.data_ptr(), X0d.data_ptr(), N, ind, 0.1, R, pos, out.data_ptr())
        f, m = host(out)
        fr, mr = O.tip_force(X0 + u, ind, 0.1, R, pos)
        assert f == pytest.approx(fr, rel=1e-12) and m == pytest.approx(mr, rel=1e-12)


@pytest.mark.parametrize("law", [(1.0, 0.0, 1.0, 0.0005), (1.0, 0.0, 1.0, 0.0), (2.5, 0.3, 0.01, 0.0),
                                 (0.7, 1.2, 0.05, 0.0003)])
def test_viscoplastic_kernel_parity(ctx, law):
    """Flow exponent m, hardening exponent n, hardening threshold eps0 and yield offset: the linear
    law with and without the offset (its own code path) and two power laws."""
    import torch

    from nofluidx3d_amd import _lib

    P, T = small_mesh()
    N, nT = len(P), len(T)
    dminv, v0, ev, vol = gpu_precompute(ctx, P, T)
    x = deformed(P, amp=0.15)
    x32 = O.soa(x.astype(np.float32).astype(np.float64))
    E, nu, ratio = 0.01, 0.46, 0.75
    mu = E / (2 * (1 + nu))
    m, n, eps0, sy = law
    prm = dict(mu1=ratio * mu, mu2=(1 - ratio) * mu, kappa=E / (3 * (1 - 2 * nu)), m=m, Y=0.002, rate=5e-3,
               eps0=eps0, n=n, sy=sy)
    vp_o = O.VPParams(*prm.values())
    vp_g = _lib.Viscoplastic(*prm.values())
    Fp_r = O.soa(np.tile(np.eye(3).reshape(1, 9), (nT, 1)))
    acc_r = np.zeros(nT)
    Fp_d = dev(Fp_r, np.float32)
    acc_d = torch.zeros(nT, dtype=torch.float32, device="cuda:0")
    vm_d = torch.zeros(nT, dtype=torch.float32, device="cuda:0")
    pr_d = torch.zeros(nT, dtype=torch.float32, device="cuda:0")
    xd, td = dev(x32, np.float32), dev(O.soa(T, np.uint32))
    for it in range(5):
        f_r, vm_r, pr_r = np.zeros(3 * N), np.zeros(nT), np.zeros(nT)
        O.viscoplastic(vp_o, x32, O.soa(T, np.uint32), O.soa(ev), np.ascontiguousarray(vol), f_r, acc_r, Fp_r, vm_r,
                       pr_r)
        f_d = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
        call("nfx_viscoplastic_force", ctx, vp_g, xd.data_ptr(), td.data_ptr(), dminv.data_ptr(), N, nT,
             Fp_d.data_ptr(), acc_d.data_ptr(), vm_d.data_ptr(), pr_d.data_ptr(), f_d.data_ptr(), None)
        assert np.max(np.abs(host(f_d) - f_r)) < 3e-5 * np.max(np.abs(f_r)), it
        assert np.max(np.abs(host(vm_d) - vm_r)) < 3e-5 * np.max(vm_r)
        assert np.max(np.abs(host(pr_d) - pr_r)) < 3e-5 * np.max(np.abs(pr_r))
        assert np.max(np.abs(host(acc_d) - acc_r)) < 3e-5 * np.max(acc_r) + 1e-12
        assert np.max(np.abs(host(Fp_d) - Fp_r)) < 1e-6
    assert acc_r.max() > 0  # flow happened


def test_vp_deviation_conversions(ctx):
    """Q = Fp^-1 - I (the fused viscoplastic state) from Fp and back, against FP64 numpy on random
    plastic deformations; and the position forms x = X0 + u, u = x - X0."""
    import torch

    nT = 5000
    Fp = (np.eye(3) + 0.05 * RNG.normal(size=(nT, 3, 3))).astype(np.float32).astype(np.float64)
    Fpd = dev(Fp.reshape(nT, 9).T.reshape(-1), np.float32)
    Q = torch.empty(9 * nT, dtype=torch.float32, device="cuda:0")
    call("nfx_vp_q_from_fp", ctx, Fpd.data_ptr(), nT, Q.data_ptr())
    Qg = host(Q).reshape(9, -1).T.reshape(-1, 3, 3)
    Qref = np.linalg.inv(Fp) - np.eye(3)
    # an FP32 Fp near I: the inverse's diagonal carries ~ulp(1) = 1.2e-7 before I is subtracted
    assert np.max(np.abs(Qg - Qref)) < 5e-7
    back = torch.empty_like(Q)
    call("nfx_vp_fp_from_q", ctx, Q.data_ptr(), nT, back.data_ptr())
    Fb = host(back).reshape(9, -1).T.reshape(-1, 3, 3)
    assert np.max(np.abs(Fb - Fp)) < 5e-7
    N = 4000
    X0 = RNG.normal(scale=40.0, size=3 * N).astype(np.float32)
    u = RNG.normal(scale=0.5, size=3 * N).astype(np.float32)
    X0d, ud = dev(X0), dev(u)
    x = torch.empty_like(X0d)
    call("nfx_points_from_disp", ctx, X0d.data_ptr(), ud.data_ptr(), N, x.data_ptr())
    assert np.array_equal(host(x, np.float32), X0 + u)
    u2 = torch.empty_like(X0d)
    call("nfx_disp_from_points", ctx, X0d.data_ptr(), x.data_ptr(), N, u2.data_ptr())
    assert np.array_equal(host(u2, np.float32), (X0 + u) - X0)


def test_graph_is_captured_once_per_plan(ctx):
    """Whole 100-step blocks replay one captured graph whatever the call's step count; the rest of
    a call runs eagerly (a second call with the same count, or another count, captures nothing)."""
    from nofluidx3d_amd import _lib

    def captures():
        n = ctypes.c_int64()
        _lib.call("nfx_graph_captures", ctx, ctypes.byref(n))
        return n.value

    sim = build_sim(scatter="tiled")
    c0 = captures()
    sim.advance(20)
    assert captures() == c0  # shorter than a block: eager
    sim.advance(250)
    c1 = captures()
    assert c1 == c0 + 1
    sim.advance(250)
    sim.advance(130)
    sim.advance(20)
    assert captures() == c1


# ------------------------------------------------------------------------------------------------
# fused multi-step runs through the Simulation API
# ------------------------------------------------------------------------------------------------


LEGACY_SETTERS = {"LE": "setInteractionLinearElastic", "LD": "setInteractionLinearPoroElastic",
                  "NH2": "setInteractionSecondOrderNeoHookean", "PNH": "setInteractionPoroNeoHookean"}
# the Tetra plugin's other laws (config CELL keys)
ELASTIC_LAWS = {"SVK": {"model": "SAINT_VENANT_KIRCHHOFF"}, "SVKK": {"model": "SAINT_VENANT_KIRCHHOFF_KELLNBERGER"},
                "MR": {"model": "MOONEY_RIVLIN", "MooneyRivlinRatio": 0.75}}


def with_poles(P, T):
    """The mesh relabelled so that nodes 3 and 4 are its top and bottom poles on the y axis (moved
    onto x = z = 0 from a few hundredths of the lattice spacing), the nodes fixTopBottom pins."""
    r = np.hypot(P[:, 0], P[:, 2])
    near = np.flatnonzero((r < 0.2) & (np.arange(len(P)) >= 2))
    top, bottom = near[np.argmax(P[near, 1])], near[np.argmin(P[near, 1])]
    assert len({top, bottom, 3, 4}) == 4
    m = np.arange(len(P))
    m[[3, top]] = m[[top, 3]]
    m[[4, bottom]] = m[[bottom, 4]]
    P = P[m].copy()
    P[3:5, [0, 2]] = 0.0
    inv = np.argsort(m)
    return P, inv[T]


def build_sim(cfg_overrides=None, scatter="atomic", fused=True, use_graph=True, mesh=(4.0, 0.8, 1), material="NH",
              indenter="sphere", steps=400, compact=False, yield_si=5.0, node_fused=True, fix=False):
    from nofluidx3d_amd import configs
    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    cfg = configs.small_cell(outdir="/tmp/nfx_test/", steps=steps, frames=4)
    cfg["MESH"] = {"radius_um": mesh[0], "spacing_um": mesh[1], "seed": mesh[2]}
    cfg["CELL"]["RadiusSI"] = mesh[0] * 1e-6
    cfg["CELL"]["RadiusSIM"] = 2.5 * mesh[0]
    cfg["SphereRadiusSI"] = 0.4 * mesh[0] * 1e-6
    cfg["InitialDistance"] = -0.5
    cfg.update(cfg_overrides or {})
    if material == "VP":
        cfg["CELL"].update({"MooneyRivlinRatio": 0.75, "model": "MOONEY_RIVLIN", "ViscositySI": 10.0,
                            "YieldStressSI": yield_si})
    cfg["CELL"].update(ELASTIC_LAWS.get(material, {}))
    P, T = configs.make_mesh(cfg)
    if fix:
        P, T = with_poles(P, T)
    tile_k = {"tiled256": 256, "tiled128": 128}.get(scatter, 512)
    scatter = "tiled" if scatter.startswith("tiled") else scatter
    sim = Simulation(cfg, scatter=scatter, fused=fused, use_graph=use_graph, mesh=vtk.from_tetra(P, T),
                     verbose=False, tile_k=tile_k, fuse_node_pass=node_fused)
    sim.setInteractionSubstrate()
    if indenter == "sphere":
        sim.setInteractionSphere()
    else:
        sim.setInteractionPlane()
    if material == "VP":
        sim.setInteractionFiniteStrainViscoplastic(compactState=compact)
    elif material in LEGACY_SETTERS:
        cfg["CELL"].update({"PoroElasticFactor": 0.3, "SecondOrderNeoHookeanFactor": 0.5, "PorousVolumeFraction": 0.4})
        getattr(sim, LEGACY_SETTERS[material])()
    else:
        sim.setInteractionTetra()
    sim.setInteractionVelocityVerlet(fixTopBottom=fix)
    return sim


def oracle_for(sim, nthreads=1):
    from nofluidx3d_amd.interactions import FiniteStrainViscoplastic, LegacyElastic, Sphere, Substrate, Tetra

    prm = O.RunParams()
    prm.nthreads = nthreads
    for inter in sim.interactions:
        if isinstance(inter, LegacyElastic):
            p = inter.params
            prm.model = inter.model
            prm.legacy = O.LegacyParams(E=p.E, nu=p.nu, shear=p.shear, bulk=p.bulk, shear2=p.shear2,
                                        phi=p.volume_fraction)
        elif isinstance(inter, Tetra):
            prm.model = O.MR if inter.model_code() == 2 else inter.model_code()
            prm.E, prm.nu, prm.ratio = (float(v) for v in (inter.youngsModulus, inter.poissonRatio,
                                                             inter.mooneyRivlinRatio))
        elif isinstance(inter, FiniteStrainViscoplastic):
            prm.model = O.VISCOPLASTIC
            p = inter.params
            prm.vp = O.VPParams(p.mu1, p.mu2, p.kappa, p.flow_exponent, p.yield_stress, p.flow_rate,
                                p.hardening_threshold, p.hardening_exponent, p.yield_offset)
        elif isinstance(inter, Substrate):
            prm.substrate, prm.substrate_k, prm.wall = 1, inter.forceConst, inter.wallPosition
        elif isinstance(inter, Sphere):
            prm.indenter, prm.ind_k, prm.ind_radius = 1, inter.forceConst, inter.radius
            prm.ind_start, prm.ind_speed, prm.ind_move_steps = inter.sphereFunc.law
        elif hasattr(inter, "wallFunc"):
            prm.indenter, prm.ind_k, prm.ind_radius = 2, inter.forceConst, 0.0
            prm.ind_start, prm.ind_speed, prm.ind_move_steps = inter.wallFunc.law
        elif hasattr(inter, "fixTopBottom"):
            prm.fix_top_bottom = int(inter.fixTopBottom)
    return O.OracleState(sim.cell.points, sim.cell.tetras, prm)


def run_gpu(sim, nsteps):
    import torch

    hist = torch.zeros(nsteps, dtype=torch.float64, device="cuda:0")
    sim.advance(nsteps, tip_hist=hist)
    return sim.positions(), host(hist)


def assert_parity(sim, gx, gf, ox, of):
    R = sim.parameters["CELL"]["RadiusSIM"]
    pos_err = np.max(np.abs(gx - ox)) / R
    f_err = np.max(np.abs(gf - of)) / np.max(np.abs(of))
    assert np.max(np.abs(of)) > 0
    assert pos_err < 1e-4, pos_err
    assert f_err < 1e-4, f_err
    return pos_err, f_err


@pytest.mark.parametrize("scatter", ["atomic", "csr", "tiled", "tiled256", "tiled128"])
@pytest.mark.parametrize("indenter", ["sphere", "plane"])
def test_fused_run_matches_oracle(ctx, scatter, indenter):
    sim = build_sim(scatter=scatter, indenter=indenter)
    if scatter == "tiled128":
        assert sim.plan().tile_k == 128 and sim.scatter == "tiled"
    gx, gf = run_gpu(sim, 600)
    ref = oracle_for(sim)
    of = ref.run(600)
    assert_parity(sim, gx, gf, ref.positions(), of)


@pytest.mark.parametrize("node_fused", [True, False], ids=["one_launch", "two_launches"])
@pytest.mark.parametrize("material", ["NH", "VP"])
def test_fused_calls_of_every_length_match_oracle(ctx, material, node_fused):
    """A multi-step call carries (u, w) with w = a*v + f_old/2 between its first and last step; a
    one-step call keeps (v, f_old). Calls of 1, 2, 3 and more steps (eager and graph-replayed), in
    any order, hand (v, f_old) to the next call, as the reference's per-step loop does; with the
    node-fused step each call also begins and ends in the plan's own buffers."""
    import torch

    sim = build_sim(material=material, scatter="tiled", node_fused=node_fused)
    assert bool(sim.plan().disp_b) == node_fused
    vv = sim.interactions[-1]
    chunks = [1, 2, 1, 3, 7, 1, 150, 2, 1, 32]
    hist = torch.zeros(sum(chunks), dtype=torch.float64, device="cuda:0")
    ref = oracle_for(sim)
    t, of = 0, []
    for n in chunks:
        sim.advance(n, tip_hist=hist[t:t + n])
        of.append(ref.run(n).copy())
        t += n
        gv = vv.velocities()
        gfo = vv.forceOldB.cpu().numpy().astype(np.float64)
        vs = max(np.max(np.abs(ref.vel)), 1e-30)
        fs = max(np.max(np.abs(ref.fold)), 1e-30)
        assert np.max(np.abs(gv - ref.vel)) / vs < 1e-3, (n, t)
        assert np.max(np.abs(gfo - ref.fold)) / fs < 1e-3, (n, t)
    assert_parity(sim, sim.positions(), host(hist), ref.positions(), np.concatenate(of))


@pytest.mark.parametrize("fix", [False, True], ids=["free", "fix_top_bottom"])
@pytest.mark.parametrize("material", ["NH", "VP"])
def test_node_fused_step_matches_two_launch_step(ctx, material, fix):
    """The node-fused step (each node update carried by the next force pass; every tile that
    touches a node computes it, its first tile stores it) against the same step as two launches:
    the same recurrence in the same operation order, equal to within FP32 rounding of the compiled
    forms, and both at the oracle's 1e-4."""
    runs = {}
    for nf in (True, False):
        sim = build_sim(material=material, scatter="tiled256", node_fused=nf, fix=fix)
        assert bool(sim.plan().disp_b) == nf
        runs[nf] = (sim,) + run_gpu(sim, 600)
    sim, gx, gf = runs[True]
    _, ux, uf = runs[False]
    R = sim.parameters["CELL"]["RadiusSIM"]
    assert np.max(np.abs(gx - ux)) / R < 1e-6
    assert np.max(np.abs(gf - uf)) / np.max(np.abs(uf)) < 1e-6
    if fix:
        assert np.all(gx[3:5, [0, 2]] == 0.0)  # x = z = 0 at nodes 3 and 4 (VelocityVerlet.cl:17-22)
    ref = oracle_for(sim)
    of = ref.run(600)
    assert_parity(sim, gx, gf, ref.positions(), of)


@pytest.mark.parametrize("scatter", ["csr", "tiled"])
def test_viscoplastic_fused_run_matches_oracle(ctx, scatter):
    sim = build_sim(material="VP", scatter=scatter)
    gx, gf = run_gpu(sim, 600)
    ref = oracle_for(sim)
    of = ref.run(600)
    assert ref.acc.max() > 0
    assert_parity(sim, gx, gf, ref.positions(), of)


@pytest.mark.parametrize("scatter", ["csr", "tiled"])
def test_viscoplastic_state_matches_oracle(ctx, scatter):
    """Per-tetra state after 600 steps: Fp (converted from the fused step's plastic deviation Q) and
    the accumulated plastic strain, against the FP64 oracle's Fp and epsbar."""
    from nofluidx3d_amd.interactions import FiniteStrainViscoplastic

    sim = build_sim(material="VP", scatter=scatter)
    gx, gf = run_gpu(sim, 600)
    ref = oracle_for(sim)
    of = ref.run(600)
    vp = next(i for i in sim.interactions if isinstance(i, FiniteStrainViscoplastic))
    Fp = host(vp.plastic_deformation())
    acc = host(vp.accumulatedStrainB)
    assert np.max(np.abs(Fp - np.eye(3).reshape(9).repeat(len(acc)))) > 1e-4  # plastic flow happened
    assert np.max(np.abs(Fp - ref.Fp)) < 1e-4
    assert np.max(np.abs(acc - ref.acc)) < 1e-3 * np.max(ref.acc)
    assert_parity(sim, gx, gf, ref.positions(), of)


def test_viscoplastic_switching_between_fused_and_per_plugin_steps(ctx):
    """Tiled fused steps (displacements, plastic deviation Q), then per-plugin steps (absolute
    positions, Fp), then fused again: the state follows each switch and the run matches the
    oracle."""
    import torch

    sim = build_sim(material="VP", scatter="tiled")
    hist = torch.zeros(600, dtype=torch.float64, device="cuda:0")
    sim.advance(300, tip_hist=hist[:300])
    sim.fused = False
    for _ in range(30):
        sim.timeStep()
    sim.fused = True
    sim.advance(270, tip_hist=hist[330:])
    ref = oracle_for(sim)
    of = ref.run(600)
    gf = host(hist)
    keep = np.r_[0:300, 330:600]
    assert_parity(sim, sim.positions(), gf[keep], ref.positions(), of[keep])


@pytest.mark.parametrize("law", [4, 5, 6, 7])
def test_legacy_law_kernel_parity(ctx, law):
    """nfx_legacy_force (atomic scatter and per-corner records + CSR gather) against the oracle's
    restatement of the legacy kernel: nodal forces, von Mises stress and pressure."""
    import torch

    from nofluidx3d_amd import _lib

    P, T = small_mesh()
    N, nT = len(P), len(T)
    dminv, v0, ev, vol = gpu_precompute(ctx, P, T)
    x = deformed(P, amp=0.12)
    x32 = O.soa(x.astype(np.float32).astype(np.float64))
    E, nu = 0.01, 0.42
    mu, K = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
    prm = dict(E=E, nu=nu, shear=mu, bulk=K, shear2=0.5 * mu, phi=0.4)
    f_r, vm_r, pr_r = np.zeros(3 * N), np.zeros(nT), np.zeros(nT)
    O.legacy_force(law - 4, O.LegacyParams(**prm), x32, O.soa(T, np.uint32), O.soa(ev), f_r, vm_r, pr_r)
    mat = _lib.Elastic(law, prm["E"], prm["nu"], 1.0, prm["shear"], prm["bulk"], prm["shear2"], prm["phi"])
    xd, td = dev(x32, np.float32), dev(O.soa(T, np.uint32))
    f_d = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    vm_d = torch.zeros(nT, dtype=torch.float32, device="cuda:0")
    pr_d = torch.zeros(nT, dtype=torch.float32, device="cuda:0")
    call("nfx_legacy_force", ctx, mat, xd.data_ptr(), td.data_ptr(), dminv.data_ptr(), N, nT, vm_d.data_ptr(),
         pr_d.data_ptr(), f_d.data_ptr(), None)
    assert np.max(np.abs(host(f_d) - f_r)) < 3e-5 * np.max(np.abs(f_r))
    assert np.max(np.abs(host(vm_d) - vm_r)) < 3e-5 * np.max(vm_r)
    assert np.max(np.abs(host(pr_d) - pr_r)) < 3e-5 * np.max(np.abs(pr_r))
    off, slots = csr(T, N)
    corners = torch.zeros(12 * nT, dtype=torch.float32, device="cuda:0")
    g_d = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    call("nfx_legacy_force", ctx, mat, xd.data_ptr(), td.data_ptr(), dminv.data_ptr(), N, nT, None, None, None,
         corners.data_ptr())
    call("nfx_gather_corners", ctx, off.data_ptr(), slots.data_ptr(), corners.data_ptr(), N, g_d.data_ptr())
    assert np.max(np.abs(host(g_d) - f_r)) < 3e-5 * np.max(np.abs(f_r))


@pytest.mark.parametrize("material", ["LE", "LD", "NH2", "PNH"])
@pytest.mark.parametrize("scatter", ["atomic", "tiled"])
def test_legacy_law_fused_run_matches_oracle(ctx, material, scatter):
    """200 steps with the tip starting 0.2 above the cell and the wall 0.3 below it: the linear
    laws are not objective and, as in the reference, fail under the deep, fast indentation of the
    other fused tests (the oracle itself diverges there after ~250 steps)."""
    sim = build_sim(material=material, scatter=scatter, cfg_overrides={"InitialDistance": 0.2, "SubstrateDistance": 0.3})
    gx, gf = run_gpu(sim, 200)
    ref = oracle_for(sim)
    of = ref.run(200)
    assert_parity(sim, gx, gf, ref.positions(), of)


@pytest.mark.parametrize("node_fused", [True, False], ids=["one_launch", "two_launches"])
@pytest.mark.parametrize("material", ["SVK", "SVKK", "MR"])
def test_elastic_laws_tiled_run_matches_oracle(ctx, material, node_fused):
    """The Tetra plugin's SVK, SVKK and MR (ratio 0.75) laws through the tiled step in both step
    forms, 300 steps with the tip starting 0.2 above the cell (the SVK laws soften in compression and
    are run short of the deep indentation)."""
    sim = build_sim(material=material, scatter="tiled256", node_fused=node_fused,
                    cfg_overrides={"InitialDistance": 0.2, "SubstrateDistance": 0.3})
    assert sim.plan().model == {"SVK": O.SVK, "SVKK": O.SVKK, "MR": O.MR}[material]
    assert bool(sim.plan().disp_b) == node_fused
    gx, gf = run_gpu(sim, 300)
    ref = oracle_for(sim)
    of = ref.run(300)
    assert_parity(sim, gx, gf, ref.positions(), of)


def test_fused_run_writes_stress_of_the_last_step(ctx):
    """With write_stress on, the last fused step of a call writes vonMises and pressure: equal to
    the oracle's law evaluated at the state that step started from."""
    from nofluidx3d_amd.interactions import LegacyElastic

    sim = build_sim(material="NH2", scatter="tiled")
    inter = next(i for i in sim.interactions if isinstance(i, LegacyElastic))
    inter.write_stress = True
    sim._plan = None
    run_gpu(sim, 300)
    ref = oracle_for(sim)
    ref.run(299)
    p = inter.params
    f = np.zeros_like(ref.points)
    vm, pr = np.zeros(ref.T), np.zeros(ref.T)
    O.legacy_force(2, O.LegacyParams(E=p.E, nu=p.nu, shear=p.shear, bulk=p.bulk, shear2=p.shear2,
                                     phi=p.volume_fraction), ref.points, ref.tetras, ref.ev, f, vm, pr)
    gvm, gpr = host(sim.vonMises.buf), host(sim.pressure.buf)
    assert np.max(vm) > 0
    assert np.max(np.abs(gvm - vm)) < 1e-3 * np.max(vm)
    assert np.max(np.abs(gpr - pr)) < 1e-3 * np.max(np.abs(pr))


@pytest.mark.parametrize("scatter", ["csr", "tiled", "tiled_two_launches"])
def test_deterministic_modes_are_bitwise_reproducible_and_graph_equals_eager(ctx, scatter):
    nf = scatter == "tiled"  # the tiled step in the node-fused form, or as two launches per step
    scatter = "tiled" if scatter.startswith("tiled") else scatter
    a = build_sim(scatter=scatter, use_graph=True, node_fused=nf)
    xa, fa = run_gpu(a, 250)
    b = build_sim(scatter=scatter, use_graph=True, node_fused=nf)
    xb, fb = run_gpu(b, 250)
    c = build_sim(scatter=scatter, use_graph=False, node_fused=nf)
    xc, fc = run_gpu(c, 250)
    assert np.array_equal(xa, xb) and np.array_equal(fa, fb)
    assert np.array_equal(xa, xc) and np.array_equal(fa, fc)


def test_per_plugin_path_matches_fused(ctx):
    """timeStep() plugin by plugin (the reference's loop) vs the fused device step."""
    a = build_sim(scatter="csr", fused=False)
    for _ in range(60):
        a.timeStep()
    b = build_sim(scatter="csr", fused=True)
    b.advance(60)
    xa, xb = a.positions(), b.positions()
    assert np.max(np.abs(xa - xb)) < 1e-5 * a.parameters["CELL"]["RadiusSIM"]


def test_simulation_run_writes_frames_and_simdata(ctx, tmp_path):
    from nofluidx3d_amd.util import vtk

    sim = build_sim(cfg_overrides={"OutputDir": str(tmp_path) + "/"}, steps=80)
    sim.parameters["VTKFramesTotal"] = 4
    sim.run()
    data = np.loadtxt(str(tmp_path / "simdata.dat"))
    assert data.shape == (5, 3)
    assert np.array_equal(data[:, 0], [0, 20, 40, 60, 80])
    # the recorded force equals the oracle's indenter force at the same times
    ref = oracle_for(build_sim(steps=80))
    of = ref.run(81)
    scale = sim.p0 * sim.L0**2
    assert np.max(np.abs(data[:, 2] / scale - of[[0, 20, 40, 60, 80]])) < 1e-4 * np.max(np.abs(of))
    g = vtk.read(str(tmp_path / "vtkfiles" / "frame80.vtk"))
    assert np.allclose(g.points, sim.positions())


def test_small_cell_full_run_matches_oracle(ctx):
    """The plain small cell (R 8 um, 0.8 um lattice, seed 1; ~25k tetra) over its full 20 000
    steps: positions and the 201-sample tip-force curve against the FP64 oracle."""
    import torch

    from nofluidx3d_amd import configs
    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    cfg = configs.small_cell(outdir="/tmp/nfx_c1/")
    P, T = configs.make_mesh(cfg)
    sim = Simulation(cfg, scatter="tiled", mesh=vtk.from_tetra(P, T), verbose=False)
    sim.setInteractionSubstrate()
    sim.setInteractionSphere()
    sim.setInteractionTetra()
    sim.setInteractionVelocityVerlet()
    n = sim.Nsteps
    hist = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    sim.advance(n, tip_hist=hist)
    gx, gf = sim.positions(), host(hist)
    ref = oracle_for(sim, nthreads=O.max_threads())
    of = ref.run(n)
    pe, fe = assert_parity(sim, gx, gf[::100], ref.positions(), of[::100])
    print(f"small cell 20k steps: pos_err {pe:.2e} force_err {fe:.2e} max F {np.max(np.abs(of)):.3e}")


@pytest.mark.parametrize("material", ["NH", "VP"])
def test_force_pass_cache_policy_changes_no_bit(ctx, material):
    """The tiled force pass with the nontemporal cache policy gives the same bits as the default."""
    runs = []
    for nt in (0, 1):
        sim = build_sim(material=material, scatter="tiled")
        sim.stream_nt = nt
        runs.append(run_gpu(sim, 300))
        assert sim.plan().stream_nt == nt
    (xa, fa), (xb, fb) = runs
    assert np.array_equal(xa, xb) and np.array_equal(fa, fb)


@pytest.mark.parametrize("material", ["NH", "VP", "LE"])
def test_shape_from_reference_positions_matches_streamed_shape(ctx, material):
    """The tiled force pass forming Dm^-1 and V0 from the FP32 reference positions and the one
    reading the FP64-precomputed Dm^-1 and V0 both meet the oracle, and stay within 1e-5 of each
    other (the two shapes differ by FP32 rounding, ~1e-7 relative).  The linear law runs the shallow
    window of the legacy-law test (it is not objective and diverges under deep indentation)."""
    over, n = ({"InitialDistance": 0.2, "SubstrateDistance": 0.3}, 200) if material == "LE" else (None, 600)
    runs = []
    for geo in (True, False):
        sim = build_sim(material=material, scatter="tiled", cfg_overrides=over)
        sim.shape_from_ref = geo
        runs.append(run_gpu(sim, n))
        assert sim.plan().shape_from_ref == int(geo)
    ref = oracle_for(sim)
    of = ref.run(n)
    for gx, gf in runs:
        assert_parity(sim, gx, gf, ref.positions(), of)
    (xa, fa), (xb, fb) = runs
    R = sim.parameters["CELL"]["RadiusSIM"]
    assert np.max(np.abs(xa - xb)) / R < 1e-5
    assert np.max(np.abs(fa - fb)) / np.max(np.abs(fb)) < 1e-5


def test_compact_viscoplastic_state_matches_oracle_and_full_state(ctx):
    """The compact state (symmetric deviation Gd, no plastic strain for the Maxwell-type law) on the
    fused tiled step: within the oracle's bars and within 1e-5 of the full state (Q and epsbar)."""
    full = build_sim(material="VP", scatter="tiled")
    xf, ff = run_gpu(full, 600)
    sim = build_sim(material="VP", scatter="tiled", compact=True)
    gx, gf = run_gpu(sim, 600)
    p = sim.plan()
    assert p.vp_sym == 1 and not p.epsbar and sim.interactions[2].plasticDeviationSymB.numel() == 6 * sim.numTetra
    ref = oracle_for(sim)
    of = ref.run(600)
    assert_parity(sim, gx, gf, ref.positions(), of)
    R = sim.parameters["CELL"]["RadiusSIM"]
    assert np.max(np.abs(gx - xf)) / R < 1e-5
    assert np.max(np.abs(gf - ff)) / np.max(np.abs(ff)) < 1e-5
    # Fp is not kept: the per-plugin step and the Fp read-back refuse
    with pytest.raises(RuntimeError, match="compact"):
        sim.interactions[2].plastic_deformation()
    eager = build_sim(material="VP", scatter="tiled", compact=True, fused=False)
    with pytest.raises(RuntimeError, match="compact"):
        eager.timeStep()


def test_plastic_activity_flags_change_no_bit(ctx):
    """The tiled viscoplastic step with per-tile plastic-activity flags (tiles whose tetra never
    flowed read and write no state) against the same step without flags: bitwise equal, with both
    kinds of tile present; the flags the kernel kept equal those recomputed from the state."""
    import torch

    from nofluidx3d_amd import _lib, openCL

    runs = []
    for flags in (True, False):
        sim = build_sim(material="VP", scatter="tiled256", mesh=(6.0, 0.6, 1), yield_si=60.0)
        sim.use_plastic_flags = flags
        runs.append(run_gpu(sim, 600))
        assert bool(sim.plan().vp_active) == flags
        if flags:
            kept = host(sim._vp_active, np.int64)
            _lib.call("nfx_vp_tile_flags", openCL.handle(), sim.plan())
            assert np.array_equal(host(sim._vp_active, np.int64), kept)
            print(f"{int(kept.sum())} of {len(kept)} tiles have flowed")
            assert 0 < kept.sum() < len(kept), kept
            q = host(sim.interactions[2].plasticDeviationB).reshape(9, -1)
            assert np.any(q != 0)
    (xa, fa), (xb, fb) = runs
    assert np.array_equal(xa, xb) and np.array_equal(fa, fb)
    torch.cuda.synchronize()


def test_checkpoint_restart_is_bitwise(ctx, tmp_path):
    """save_state after 300 steps, then 200 more steps, against a fresh Simulation that loads the
    checkpoint and runs the same 200: bitwise equal positions, tip force and plastic state (the
    restored run recomputes its plastic-activity flags)."""
    import torch

    a = build_sim(material="VP", scatter="tiled")
    run_gpu(a, 300)
    path = str(tmp_path / "state.npz")
    a.save_state(path)
    xa, fa = run_gpu(a, 200)
    b = build_sim(material="VP", scatter="tiled")
    b.load_state(path)
    assert b.time == 300
    xb, fb = run_gpu(b, 200)
    assert np.array_equal(xa, xb) and np.array_equal(fa, fb)
    qa, qb = (host(s.interactions[2].plastic_deformation()) for s in (a, b))
    assert np.array_equal(qa, qb)
    c = build_sim(material="NH", scatter="tiled")
    with pytest.raises(ValueError):
        c.load_state(path)
    torch.cuda.synchronize()


def test_device_selection_before_any_torch_use(ctx):
    """A fresh process whose first GPU call is initializeOpenCLObjects (no torch use before it):
    libnfx must share torch's HIP runtime, or torch finds no device afterwards."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("from nofluidx3d_amd import openCL\n"
            "openCL.initializeOpenCLObjects(0)\n"
            "import torch\n"
            "x = torch.ones(4, device='cuda:0')\n"
            "print('ok', float(x.sum()), openCL.getDevice().name)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith("ok 4.0"), r.stdout
