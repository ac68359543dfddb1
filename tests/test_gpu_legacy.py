"""GPU parity of the legacy contacts and integrators (libnfx entry points) against the FP64 oracle's
restatements on a deformed seeded mesh, FP32 state: forces within 3e-5 relative to their maximum
(float-atomic sums in arbitrary order), positions of the integrators within 1e-6 relative."""

import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)


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


def host(t):
    import torch

    from nofluidx3d_amd import _lib, openCL

    _lib.call("nfx_stream_sync", openCL.handle())
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


def call(name, *args):
    from nofluidx3d_amd import _lib

    _lib.call(name, *args)


@pytest.fixture(scope="module")
def mesh():
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, T = generate_sphere_mesh(3.0, 0.6, 5)
    P = P * 3.0
    x = P + 0.05 * RNG.normal(size=P.shape)
    x32 = O.soa(x.astype(np.float32).astype(np.float64))
    return P, T, x32


def rel_close(g, r, tol=3e-5):
    assert np.max(np.abs(g - r)) <= tol * max(np.max(np.abs(r)), 1e-30), np.max(np.abs(g - r)) / np.max(np.abs(r))


def test_node_contacts(ctx, mesh):
    import torch

    P, T, x = mesh
    N = len(P)
    xd = dev(x, np.float32)
    # round tip: the tip centre inside the cell, so both the sphere and the cylinder act
    fr = np.zeros(3 * N)
    O.round_tip(fr, x, 3.0, 2.0)
    fd = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    call("nfx_round_tip_force", ctx, fd.data_ptr(), xd.data_ptr(), N, 3.0, 2.0)
    assert np.abs(fr).max() > 0
    rel_close(host(fd), fr)
    # adhesive plane on every other point
    on = (np.arange(N) % 2).astype(np.int32)
    wall = x[N:2 * N].min() - 0.5
    fr = np.zeros(3 * N)
    O.adhesive_plane(fr, x, wall, 0.01, 1.0, on)
    fd.zero_()
    ond = dev(on)  # keep every device buffer referenced until the kernels have run
    call("nfx_adhesive_plane_force", ctx, fd.data_ptr(), xd.data_ptr(), N, wall, 0.01, 1.0, ond.data_ptr())
    rel_close(host(fd), fr, 1e-4)  # (sigma/d)^12 in FP32
    # point zero and centre of mass
    fr = np.zeros(3 * N)
    O.point_zero_force(fr, x, 0.3)
    fd.zero_()
    call("nfx_point_zero_force", ctx, fd.data_ptr(), xd.data_ptr(), N, 0.3)
    rel_close(host(fd), fr, 1e-6)
    com = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    call("nfx_center_of_mass", ctx, xd.data_ptr(), N, com.data_ptr())
    cr = O.center_of_mass(x)
    assert np.allclose(host(com), cr, rtol=0, atol=1e-6 * np.abs(x).max())
    fr = np.zeros(3 * N)
    O.com_force(fr, 0.7, host(com))
    fd.zero_()
    call("nfx_com_force", ctx, fd.data_ptr(), N, 0.7, com.data_ptr())
    rel_close(host(fd), fr, 1e-6)


def test_integrators(ctx, mesh):
    import torch

    P, T, x = mesh
    N = len(P)
    f = 1e-3 * RNG.normal(size=3 * N)
    # over-damped
    xr, fr = x.copy(), f.astype(np.float32).astype(np.float64)
    O.over_damped(xr, fr, True)
    xd, fd = dev(x, np.float32), dev(f, np.float32)
    call("nfx_over_damped", ctx, xd.data_ptr(), fd.data_ptr(), N, 1)
    assert np.max(np.abs(host(xd) - xr)) < 1e-6 * np.abs(x).max()
    assert np.all(host(fd) == 0)
    # velocity Verlet 2: 20 steps under a constant force field, per-node masses
    from nofluidx3d_amd.interactions.LegacyContacts import lumped_masses

    _, vol = O.edge_vectors_volumes(P, T)
    m = lumped_masses(T, vol, N)
    xr, vr, forr = x.copy(), np.zeros(3 * N), np.zeros(3 * N)
    xd, vd, fod = dev(x, np.float32), torch.zeros(3 * N, device="cuda:0"), torch.zeros(3 * N, device="cuda:0")
    md = dev(m, np.float32)
    f32 = f.astype(np.float32).astype(np.float64)
    for _ in range(20):
        fcur = f32.copy()
        O.velocity_verlet2(xr, vr, fcur, forr, m.astype(np.float32).astype(np.float64), 0.1)
        fd = dev(f32, np.float32)
        call("nfx_velocity_verlet2", ctx, xd.data_ptr(), vd.data_ptr(), fd.data_ptr(), fod.data_ptr(), md.data_ptr(), N,
             0.1)
    assert np.max(np.abs(host(xd) - xr)) < 1e-6 * np.abs(x).max()
    assert np.max(np.abs(host(vd) - vr)) < 1e-5 * np.abs(vr).max()


@pytest.mark.parametrize("tip_shape", [0, 1])
def test_volume_integrals(ctx, mesh, tip_shape):
    import torch

    from nofluidx3d_amd import _lib

    P, T, x = mesh
    N, nT = len(P), len(T)
    top = x[N:2 * N].max()
    R, k, alpha = 2.0, 0.8, 0.35
    tip = (0.3, top + R - 1.0, -0.2)
    fr = np.zeros(3 * N)
    O.volume_quadrature(tip_shape, fr, x, O.soa(T, np.uint32), R, alpha, tip, k)
    xd, td = dev(x, np.float32), dev(O.soa(T, np.uint32))
    fd = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    call("nfx_volume_integral_force", ctx, tip_shape, xd.data_ptr(), td.data_ptr(), N, nT, R, alpha, *tip, k,
         fd.data_ptr(), None)
    rel_close(host(fd), fr)
    # per-corner records + deterministic CSR gather
    off = np.zeros(N + 1, np.uint32)
    slots = np.zeros(4 * nT, np.uint32)
    tets = O.soa(T, np.uint32)
    _lib.call("nfx_build_node_csr_host", tets.ctypes.data_as(ctypes.c_void_p), N, nT,
              off.ctypes.data_as(ctypes.c_void_p), slots.ctypes.data_as(ctypes.c_void_p))
    corners = torch.zeros(12 * nT, dtype=torch.float32, device="cuda:0")
    gd = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    call("nfx_volume_integral_force", ctx, tip_shape, xd.data_ptr(), td.data_ptr(), N, nT, R, alpha, *tip, k, None,
         corners.data_ptr())
    offd, slotsd = dev(off), dev(slots)
    call("nfx_gather_corners", ctx, offd.data_ptr(), slotsd.data_ptr(), corners.data_ptr(), N, gd.data_ptr())
    rel_close(host(gd), fr)


def test_surface_adhesion(ctx, mesh):
    import torch

    from nofluidx3d_amd.interactions.LegacyContacts import surface_triangles

    P, T, x = mesh
    N = len(P)
    tri = surface_triangles(T)
    wall = x[N:2 * N].min() - 0.4
    fr = np.zeros(3 * N)
    O.adhesive_plane_surface(fr, x, O.soa(tri, np.uint32), wall, 0.02, 1.0)
    fd = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    xd, trid = dev(x, np.float32), dev(O.soa(tri, np.uint32))
    call("nfx_adhesive_plane_surface_force", ctx, fd.data_ptr(), xd.data_ptr(), trid.data_ptr(), N, len(tri), wall,
         0.02, 1.0)
    assert np.abs(fr).max() > 0
    rel_close(host(fd), fr, 1e-4)


def test_hard_surface_adhesion(ctx, mesh):
    """InteractionHardAdhesivePlaneSurfaceIntegral.cl: exp(-k d) - A exp(-d) by the 3-point rule, against
    the oracle.  The rule samples the edge midpoints; the wall lies 0.1 above the lowest of them, so
    that the surface there (below d* = ln(A) / (1 - k) = 0.255) is repelled and the rest attracted."""
    import torch

    from nofluidx3d_amd.interactions.LegacyContacts import surface_triangles

    P, T, x = mesh
    N = len(P)
    tri = surface_triangles(T)
    y = x[N:2 * N]
    mids = np.concatenate([(y[tri[:, a]] + y[tri[:, b]]) / 2 for a, b in ((0, 2), (1, 2), (0, 1))])
    wall = mids.min() + 0.1
    A, k = 0.6, 3.0
    fr = np.zeros(3 * N)
    O.hard_adhesive_plane_surface(fr, x, O.soa(tri, np.uint32), wall, A, k)
    fd = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
    xd, trid = dev(x, np.float32), dev(O.soa(tri, np.uint32))
    call("nfx_hard_adhesive_plane_surface_force", ctx, fd.data_ptr(), xd.data_ptr(), trid.data_ptr(), N, len(tri),
         wall, A, k)
    fy = fr.reshape(3, -1)[1]
    assert fy.max() > 0 and fy.min() < 0  # repulsion near the wall, adhesion farther up
    rel_close(host(fd), fr, 1e-4)


def test_linear_viscoelastic_steps(ctx, mesh):
    """nfx_linear_viscoelastic_force over 12 steps of a changing deformation (state carried from
    step to step) against the oracle: forces, von Mises stress, pressure and both state arrays."""
    import torch

    from nofluidx3d_amd import _lib

    P, T, x0 = mesh
    N, nT = len(P), len(T)
    ev, vol = O.edge_vectors_volumes(P, T)
    evd, vold = dev(O.soa(ev)), dev(np.ascontiguousarray(vol))
    dminv = torch.empty(9 * nT, dtype=torch.float32, device="cuda:0")
    v0 = torch.empty(nT, dtype=torch.float32, device="cuda:0")
    call("nfx_precompute_f64", ctx, evd.data_ptr(), vold.data_ptr(), nT, dminv.data_ptr(), v0.data_ptr())
    w, tau = [0.3, 0.25, 0.1], [3.0, 15.0, 60.0]
    Po = O.lve_params(0.01, 0.42, w, tau)
    Pg = _lib.Viscoelastic()
    Pg.E, Pg.nu, Pg.nterms = 0.01, 0.42, 3
    for j in range(3):
        Pg.weights[j], Pg.taus[j] = w[j], tau[j]
    do_r, tr_r = np.zeros(9 * nT), np.zeros(27 * nT)
    do_d = torch.zeros(9 * nT, dtype=torch.float32, device="cuda:0")
    tr_d = torch.zeros(27 * nT, dtype=torch.float32, device="cuda:0")
    vm_d = torch.zeros(nT, dtype=torch.float32, device="cuda:0")
    pr_d = torch.zeros(nT, dtype=torch.float32, device="cuda:0")
    td = dev(O.soa(T, np.uint32))
    for s in range(12):
        # a volumetric part (trace 0.3) so that the pressure is not rounding noise around zero; strains
        # of 1-14 % so that F - I keeps most of its FP32 digits (the law is small-strain and linear)
        Fh = np.eye(3) + 0.04 * (s + 1) * np.array([[0.2, 0.5, 0.0], [0.1, -0.1, 0.2], [0.0, 0.4, 0.2]])
        x = O.soa((P @ Fh.T).astype(np.float32).astype(np.float64))
        fr, vm_r, pr_r = np.zeros(3 * N), np.zeros(nT), np.zeros(nT)
        O.linear_viscoelastic(Po, x, O.soa(T, np.uint32), O.soa(ev), fr, do_r, tr_r, vm_r, pr_r)
        xd = dev(x, np.float32)
        fd = torch.zeros(3 * N, dtype=torch.float32, device="cuda:0")
        call("nfx_linear_viscoelastic_force", ctx, Pg, xd.data_ptr(), td.data_ptr(), dminv.data_ptr(), N, nT,
             do_d.data_ptr(), tr_d.data_ptr(), vm_d.data_ptr(), pr_d.data_ptr(), fd.data_ptr(), None)
        rel_close(host(fd), fr, 1e-4)
        rel_close(host(vm_d), vm_r, 1e-4)
        # pressure = kappa tr(eps): relative to the stress scale (tr(eps) cancels to a few ulps of F)
        rel_close(host(pr_d), pr_r, 1e-4 * max(np.abs(vm_r).max() / np.abs(pr_r).max(), 1.0))
    rel_close(host(do_d), do_r, 1e-4)
    rel_close(host(tr_d), tr_r, 1e-4)


def test_hard_adhesive_setter_through_simulation(ctx):
    """setInteractionHardAdhesivePlaneNonDimensional (SimulationLegacy.py:657-692) on a Simulation: the
    wall 0.2 below the cell, A = AdhesionConstant E R / (p0 L0), k = PotentialForceConst; one enqueue
    against the oracle on the cell's surface triangles."""
    from nofluidx3d_amd import configs
    from nofluidx3d_amd.interactions.LegacyContacts import surface_triangles
    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    cfg = configs.small_cell(outdir="/tmp/nfx_test_hard/", steps=10, frames=1)
    cfg["MESH"] = {"radius_um": 3.0, "spacing_um": 0.6, "seed": 5}
    cfg["CELL"]["RadiusSI"], cfg["CELL"]["RadiusSIM"] = 3e-6, 6.0
    cfg.update({"AdhesivePlaneDistance": -0.15, "AdhesionConstant": 30.0, "PotentialForceConst": 2.5})
    P, T = configs.make_mesh(cfg)
    sim = Simulation(cfg, mesh=vtk.from_tetra(P, T), verbose=False)
    inter = sim.setInteractionHardAdhesivePlaneNonDimensional()
    f = sim._record_force(inter)
    wall = -6.0 + 0.15
    A = 30.0 * cfg["CELL"]["YoungsModulusSI"] * 3e-6 / sim.p0 / sim.L0
    assert inter.wall == wall and inter.adhesionConst == pytest.approx(A, rel=1e-14) and inter.forceConst == 2.5
    fr = np.zeros(3 * sim.numPoints)
    x = O.soa(sim.cell.points)
    O.hard_adhesive_plane_surface(fr, x, O.soa(surface_triangles(sim.cell.tetras), np.uint32), wall, A, 2.5)
    fr = fr.reshape(3, -1).T
    assert np.abs(fr[:, 1]).max() > 0 and np.all(f[:, [0, 2]] == 0)
    assert np.max(np.abs(f - fr)) < 1e-4 * np.abs(fr).max()
