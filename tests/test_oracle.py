"""Pins of the FP64 oracle against mathematics (the reference has no tests or golden vectors).

Each check is chosen so that a plausible restatement mistake (a dropped term, a wrong sign, a
transposed operand, a wrong SoA index) fails it:
- hyperelastic forces: energy consistency f = -dE/dx against an independently written strain
  energy (central differences), textbook closed forms (hydrostatic, uniaxial NH, simple-shear MR
  including the normal-stress differences), rest state, rigid motions, momentum balance;
- SVK: energy consistency with Psi = lambda/2 tr(E)^2 + mu tr(E^2); SVKK's polar root against
  scipy.linalg.polar; SVKK rest state and hydrostatic closed form of its own stress law; SVKK and
  SVK both reduce to Hooke's law at small strain;
- viscoplastic: equals the MR force when Fp = I, one-step rate known answer from the shear
  closed form, plastic incompressibility, Maxwell relaxation time;
- contacts: closed forms per node; integrator: the 3-term damped Stormer recurrence and the
  terminal velocity f/xi; tip force: brute force against a numpy sum;
- scatter: sequential vs OpenMP-atomic equality to rounding.
"""

import numpy as np
import pytest
import scipy.linalg

import oracle as O

RNG = np.random.default_rng(12345)


def one_tet(points4, tets=None):
    P = np.asarray(points4, np.float64)
    t = np.array([[0, 1, 2, 3]]) if tets is None else tets
    ev, vol = O.edge_vectors_volumes(P, t)
    return O.soa(P), O.soa(t, np.uint32), O.soa(ev), np.ascontiguousarray(vol)


REF_TET = np.array([[0.1, -0.2, 0.05], [1.3, 0.1, -0.1], [0.2, 1.1, 0.3], [-0.1, 0.25, 1.2]])


def forces_of(model, E, nu, ratio, X, x, tets=None):
    """oracle nodal forces (N,3) for current positions x, reference X."""
    _, t, ev, vol = one_tet(X, tets)
    pts = O.soa(x)
    f = np.zeros_like(pts)
    O.tetra_force(model, E, nu, ratio, pts, t, ev, vol, f)
    return f.reshape(3, -1).T


def psi_mr(F, E, nu, ratio):
    """compressible Mooney-Rivlin strain energy (isochoric invariants), written from C = F^T F."""
    mu = E / (2 * (1 + nu))
    mu1, mu2 = ratio * mu, (1 - ratio) * mu
    kappa = E / (3 * (1 - 2 * nu))
    J = np.linalg.det(F)
    C = F.T @ F
    I1 = np.trace(C)
    I2 = 0.5 * (I1**2 - np.trace(C @ C))
    return 0.5 * mu1 * (J ** (-2 / 3) * I1 - 3) + 0.5 * mu2 * (J ** (-4 / 3) * I2 - 3) + 0.5 * kappa * (J - 1) ** 2


def psi_svk(F, E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    Eg = 0.5 * (F.T @ F - np.eye(3))
    return 0.5 * lam * np.trace(Eg) ** 2 + mu * np.trace(Eg @ Eg)


def energy_forces(psi, X, x, h=1e-6):
    X = np.asarray(X)
    Dm = (X[1:] - X[0]).T
    V0 = np.linalg.det(Dm) / 6
    Dinv = np.linalg.inv(Dm)

    def energy(xx):
        Ds = (xx[1:] - xx[0]).T
        return V0 * psi(Ds @ Dinv)

    f = np.zeros_like(x)
    for a in range(4):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, c] += h
            xm[a, c] -= h
            f[a, c] = -(energy(xp) - energy(xm)) / (2 * h)
    return f


def virial_stress(X, x, f):
    """sigma = -(1/v) sum_a f_a x_a^T for the nodal forces of one homogeneously strained tetra."""
    v = np.linalg.det((x[1:] - x[0]).T) / 6
    return -sum(np.outer(f[a], x[a]) for a in range(4)) / v


def deform(F, X, t=np.zeros(3)):
    return X @ F.T + t


def random_F(scale=0.15):
    return np.eye(3) + scale * RNG.uniform(-1, 1, (3, 3))


# --------------------------------------------------------------------------------------------
# precompute (TetraCell.py:26-41)
# --------------------------------------------------------------------------------------------


def test_edge_vectors_layout_and_unit_volume():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    ev, vol = O.edge_vectors_volumes(P, np.array([[0, 1, 2, 3]]))
    assert np.allclose(ev[0], np.eye(3).ravel())
    assert vol[0] == pytest.approx(1 / 6)
    P2 = P[[0, 2, 1, 3]]  # swapped orientation: signed volume flips
    assert O.edge_vectors_volumes(P2, np.array([[0, 1, 2, 3]]))[1][0] == pytest.approx(-1 / 6)
    # SoA: element k of tetra t at k*T + t (Tetra.py:53-58)
    T = 5
    ev5 = np.arange(T * 9, dtype=np.float64).reshape(T, 9)
    s = O.soa(ev5)
    for t in range(T):
        for k in range(9):
            assert s[k * T + t] == ev5[t, k]


def test_volume_sum_is_hull_volume():
    from scipy.spatial import ConvexHull

    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, T = generate_sphere_mesh(3.0, 0.6, seed=4, min_quality=0.0, morton=False)
    _, vol = O.edge_vectors_volumes(P, T)
    assert np.all(vol > 0)
    assert vol.sum() == pytest.approx(ConvexHull(P).volume, rel=1e-10)


# --------------------------------------------------------------------------------------------
# hyperelastic forces (tetra.cl)
# --------------------------------------------------------------------------------------------


@pytest.mark.parametrize("ratio", [1.0, 0.75, 0.3])
def test_mr_energy_consistency(ratio):
    E, nu = 0.01, 0.45
    for _ in range(5):
        F = random_F(0.2)
        x = deform(F, REF_TET, RNG.normal(size=3))
        x += 0.05 * RNG.normal(size=x.shape)
        f = forces_of(O.MR, E, nu, ratio, REF_TET, x)
        fe = energy_forces(lambda G: psi_mr(G, E, nu, ratio), REF_TET, x)
        assert np.max(np.abs(f - fe)) < 1e-7 * np.max(np.abs(fe))


def test_golden_homogeneous_states():
    """tests/golden/homogeneous_states.npz: textbook stresses (uniaxial Neo-Hookean, hydrostatic
    and simple-shear Mooney-Rivlin, rigidly rotated) turned into face-area nodal forces."""
    import os

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "homogeneous_states.npz"))
    for i in range(len(g["names"])):
        f = forces_of(O.MR, float(g["E"]), float(g["nu"]), float(g["ratio"][i]), g["X"], g["x"][i])
        assert np.allclose(f, g["forces"][i], rtol=1e-9, atol=1e-12 * np.abs(g["forces"][i]).max()), g["names"][i]


def test_svk_energy_consistency():
    E, nu = 0.02, 0.3
    for _ in range(5):
        x = deform(random_F(0.2), REF_TET, RNG.normal(size=3))
        f = forces_of(O.SVK, E, nu, 1.0, REF_TET, x)
        fe = energy_forces(lambda G: psi_svk(G, E, nu), REF_TET, x)
        assert np.max(np.abs(f - fe)) < 1e-7 * np.max(np.abs(fe))


@pytest.mark.parametrize("model", [O.MR, O.SVK, O.SVKK])
def test_rest_and_rigid_motion_give_zero_force(model):
    E, nu = 0.01, 0.45
    f = forces_of(model, E, nu, 0.8, REF_TET, REF_TET.copy())
    assert np.max(np.abs(f)) < 1e-15
    Q = scipy.linalg.expm(np.array([[0, -0.7, 0.2], [0.7, 0, -0.4], [-0.2, 0.4, 0]]))
    x = deform(Q, REF_TET, np.array([3.0, -1.0, 2.0]))
    f = forces_of(model, E, nu, 0.8, REF_TET, x)
    assert np.max(np.abs(f)) < 1e-12


@pytest.mark.parametrize("model", [O.MR, O.SVK])
def test_momentum_balance(model):
    x = deform(random_F(0.25), REF_TET) + 0.05 * RNG.normal(size=(4, 3))
    f = forces_of(model, 0.01, 0.4, 0.6, REF_TET, x)
    assert np.max(np.abs(f.sum(0))) < 1e-15
    torque = sum(np.cross(x[a], f[a]) for a in range(4))
    assert np.max(np.abs(torque)) < 1e-14


def test_hydrostatic_closed_form():
    E, nu = 0.01, 0.45
    kappa = E / (3 * (1 - 2 * nu))
    for ratio in (1.0, 0.5):
        for s in (0.9, 1.1):
            x = deform(s * np.eye(3), REF_TET)
            sig = virial_stress(REF_TET, x, forces_of(O.MR, E, nu, ratio, REF_TET, x))
            assert np.allclose(sig, kappa * (s**3 - 1) * np.eye(3), rtol=1e-10, atol=1e-15)


def test_uniaxial_neo_hookean_closed_form():
    E, nu = 0.01, 0.45
    mu = E / (2 * (1 + nu))
    kappa = E / (3 * (1 - 2 * nu))
    lam = 1.2
    x = deform(np.diag([lam, 1, 1]), REF_TET)
    sig = virial_stress(REF_TET, x, forces_of(O.MR, E, nu, 1.0, REF_TET, x))
    dev = mu * lam ** (-5 / 3) * (lam**2 - 1)
    assert sig[0, 0] == pytest.approx(2 / 3 * dev + kappa * (lam - 1), rel=1e-10)
    assert sig[1, 1] == pytest.approx(-1 / 3 * dev + kappa * (lam - 1), rel=1e-10)
    assert sig[2, 2] == pytest.approx(sig[1, 1], rel=1e-12)
    assert abs(sig[0, 1]) < 1e-15


def test_simple_shear_mooney_rivlin_closed_form():
    """Classic MR simple-shear results: sigma12 = (mu1+mu2) g, sigma11 - sigma22 = (mu1+mu2) g^2,
    sigma22 - sigma33 = -mu2 g^2 (Rivlin)."""
    E, nu, ratio = 0.01, 0.45, 0.6
    mu = E / (2 * (1 + nu))
    mu1, mu2 = ratio * mu, (1 - ratio) * mu
    g = 0.3
    x = deform(np.array([[1, g, 0], [0, 1, 0], [0, 0, 1.0]]), REF_TET)
    sig = virial_stress(REF_TET, x, forces_of(O.MR, E, nu, ratio, REF_TET, x))
    assert sig[0, 1] == pytest.approx((mu1 + mu2) * g, rel=1e-10)
    assert sig[0, 0] - sig[1, 1] == pytest.approx((mu1 + mu2) * g**2, rel=1e-9)
    assert sig[1, 1] - sig[2, 2] == pytest.approx(-mu2 * g**2, rel=1e-9)


def test_svkk_polar_root_and_hydrostatic():
    # approximateRoot(M) returns the symmetric polar factor P of M = U P (tetra.cl:5-12)
    import ctypes

    L = O.lib()
    if not hasattr(L, "orc_polar_stretch"):
        pytest.skip("oracle built without orc_polar_stretch")
    for _ in range(5):
        M = random_F(0.4)
        out = np.zeros(9)
        L.orc_polar_stretch(np.ascontiguousarray(M.ravel()).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert np.allclose(out.reshape(3, 3), scipy.linalg.polar(M)[1], atol=1e-12)
    # SVKK hydrostatic F = sI: S = s^2 I, rows of Force = -V0 mu (1 + (kappa(3s^2-3)-1)/s^2) s Dm^-1
    E, nu, s = 0.01, 0.3, 1.1
    mu = E / (2 * (1 + nu))
    kappa = nu / (1 - 2 * nu)
    x = deform(s * np.eye(3), REF_TET)
    f = forces_of(O.SVKK, E, nu, 1.0, REF_TET, x)
    Dm = (REF_TET[1:] - REF_TET[0]).T
    V0 = np.linalg.det(Dm) / 6
    rows = -V0 * mu * (1 + (kappa * (3 * s * s - 3) - 1) / (s * s)) * s * np.linalg.inv(Dm)
    assert np.allclose(f[1:], rows, rtol=1e-10, atol=1e-16)


def test_svkk_is_linear_elasticity_at_small_strain():
    """SVKK has no known energy to check against, but at small strain it must reduce to linear
    elasticity, as SVK does: for F = I + e H (H random, not symmetric) its forces match SVK's to
    O(e) relative, and both match Hooke's law (virial stress = lambda tr(eps) I + 2 mu eps,
    eps = e sym(H)) to O(e).  A dropped or mis-signed term of the linear part fails this."""
    E, nu = 0.01, 0.35
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    for _ in range(4):
        H = RNG.normal(size=(3, 3))
        for e in (1e-4, 1e-5):
            x = deform(np.eye(3) + e * H, REF_TET)
            fk = forces_of(O.SVKK, E, nu, 1.0, REF_TET, x)
            fs = forces_of(O.SVK, E, nu, 1.0, REF_TET, x)
            assert np.max(np.abs(fk - fs)) < 20 * e * np.max(np.abs(fs)), e
            eps = e * 0.5 * (H + H.T)
            hooke = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
            sig = virial_stress(REF_TET, x, fk)
            assert np.max(np.abs(sig - hooke)) < 20 * e * np.max(np.abs(hooke)), e


def test_sequential_and_atomic_scatter_agree():
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, T = generate_sphere_mesh(4.0, 0.8, seed=7)
    ev, vol = O.edge_vectors_volumes(P, T)
    x = P + 0.02 * RNG.normal(size=P.shape)
    fs, fp = np.zeros(3 * len(P)), np.zeros(3 * len(P))
    args = (O.MR, 0.01, 0.45, 0.8, O.soa(x), O.soa(T, np.uint32), O.soa(ev), np.ascontiguousarray(vol))
    O.tetra_force(*args, fs, nthreads=1)
    O.tetra_force(*args, fp, nthreads=4)
    assert np.allclose(fs, fp, rtol=1e-12, atol=1e-15)
    # momentum balance over the whole mesh
    assert np.max(np.abs(fs.reshape(3, -1).sum(1))) < 1e-13


# --------------------------------------------------------------------------------------------
# viscoplastic (InteractionFiniteStrainViscoplastic.cl restated with node-0 base)
# --------------------------------------------------------------------------------------------


def vp_params(E=0.01, nu=0.45, ratio=0.75, m=1.0, Y=0.002, rate=1e-3, eps0=1.0, n=0.0, sy=0.0):
    mu = E / (2 * (1 + nu))
    return O.VPParams(ratio * mu, (1 - ratio) * mu, E / (3 * (1 - 2 * nu)), m, Y, rate, eps0, n, sy)


def vp_step(P, X, x, Fp=None, acc=None, tets=None):
    _, t, ev, vol = one_tet(X, tets)
    pts = O.soa(x)
    f = np.zeros_like(pts)
    acc = np.zeros(1) if acc is None else acc
    Fp = O.soa(np.eye(3).reshape(1, 9)) if Fp is None else Fp
    vm, pr = np.zeros(1), np.zeros(1)
    O.viscoplastic(P, pts, t, ev, vol, f, acc, Fp, vm, pr)
    return f.reshape(3, -1).T, acc, Fp, vm[0], pr[0]


def test_vp_matches_mr_force_with_identity_Fp():
    E, nu, ratio = 0.01, 0.45, 0.75
    P = vp_params(E, nu, ratio, rate=0.0)
    for _ in range(5):
        x = deform(random_F(0.2), REF_TET) + 0.03 * RNG.normal(size=(4, 3))
        f_vp, acc, Fp, _, _ = vp_step(P, REF_TET, x)
        f_mr = forces_of(O.MR, E, nu, ratio, REF_TET, x)
        assert np.allclose(f_vp, f_mr, rtol=1e-9, atol=1e-15)
        assert acc[0] == 0 and np.allclose(Fp, O.soa(np.eye(3).reshape(1, 9)))


def test_vp_rest_no_flow_and_orientation_independent():
    P = vp_params()
    f, acc, Fp, vm, pr = vp_step(P, REF_TET, REF_TET.copy())
    assert np.max(np.abs(f)) < 1e-15 and acc[0] == 0 and vm == 0 and pr == 0
    # swapped node order (negative orientation): the face-area force law is orientation free
    x = deform(random_F(0.2), REF_TET)
    f1 = vp_step(P, REF_TET, x)[0]
    sw = np.array([[0, 2, 1, 3]])
    Xs, xs = REF_TET, x
    f2 = vp_step(P, Xs, xs, tets=sw)[0]
    assert np.allclose(f1, f2, rtol=1e-10, atol=1e-16)


def test_vp_one_step_rate_known_answer():
    E, nu, ratio = 0.01, 0.45, 0.6
    mu = E / (2 * (1 + nu))
    mu1, mu2 = ratio * mu, (1 - ratio) * mu
    g, m, Y, rate = 0.2, 2.0, 0.003, 1e-3
    P = vp_params(E, nu, ratio, m=m, Y=Y, rate=rate)
    x = deform(np.array([[1, g, 0], [0, 1, 0], [0, 0, 1.0]]), REF_TET)
    _, acc, Fp, vm, pr = vp_step(P, REF_TET, x)
    s12 = (mu1 + mu2) * g
    d12, d23 = (mu1 + mu2) * g**2, -mu2 * g**2
    d31 = -(d12 + d23)
    svm = np.sqrt(0.5 * (d12**2 + d23**2 + d31**2) + 3 * s12**2)
    assert vm == pytest.approx(svm, rel=1e-9)
    assert acc[0] == pytest.approx(rate * (svm / Y) ** m, rel=1e-9)
    Fpm = Fp.reshape(9, 1)[:, 0].reshape(3, 3)
    assert abs(np.linalg.det(Fpm) - 1) < 1e-3 * acc[0]


def test_vp_yield_offset_blocks_flow_below_threshold():
    P = vp_params(sy=1.0)  # threshold far above any stress here
    x = deform(np.array([[1, 0.2, 0], [0, 1, 0], [0, 0, 1.0]]), REF_TET)
    _, acc, _, vm, _ = vp_step(P, REF_TET, x)
    assert vm > 0 and acc[0] == 0


def test_vp_maxwell_relaxation_time():
    """m = 1, n = 0, held small shear: sigma12 decays like exp(-s/tau), tau = Y/(3 G eps0dot)."""
    E, nu, ratio = 0.01, 0.45, 1.0
    G = E / (2 * (1 + nu))
    Y, rate = 0.01, 2e-4
    P = vp_params(E, nu, ratio, m=1.0, Y=Y, rate=rate)
    x = deform(np.array([[1, 0.01, 0], [0, 1, 0], [0, 0, 1.0]]), REF_TET)
    acc = np.zeros(1)
    Fp = O.soa(np.eye(3).reshape(1, 9))
    vms = []
    for _ in range(400):
        _, acc, Fp, vm, _ = vp_step(P, REF_TET, x, Fp, acc)
        vms.append(vm)
    tau = Y / (3 * G * rate)
    measured = -399 / np.log(vms[-1] / vms[0])
    assert measured == pytest.approx(tau, rel=0.02)


def test_vp_shape_form_update_is_the_reference_update():
    """The tiled GPU path keeps S = Dm^-1 Fp^-1 and updates it as S' = S - r^-1 (I + Dp)^-1 Dp Fe
    with Fe = r S.  Recovering Dp from the oracle's own update (Fp' - Fp = Fe^-1 Dp F), S' must
    equal Dm^-1 Fp'^-1: the identity the GPU update rests on."""
    P = vp_params(rate=5e-2)
    I = np.eye(3)
    for _ in range(5):
        Fp0 = I + 0.05 * RNG.normal(size=(3, 3))
        x = deform(random_F(0.2), REF_TET) + 0.02 * RNG.normal(size=(4, 3))
        _, acc, Fp1, _, _ = vp_step(P, REF_TET, x, Fp=O.soa(Fp0.reshape(1, 9)).copy())
        Fp1 = Fp1.reshape(3, 3)
        Dm = (REF_TET[1:] - REF_TET[0]).T  # columns R1, R2, R3
        r = (x[1:] - x[0]).T
        F = r @ np.linalg.inv(Dm)
        Fe = F @ np.linalg.inv(Fp0)
        Dp = Fe @ (Fp1 - Fp0) @ np.linalg.inv(F)
        nd = np.abs(Dp).max()
        assert acc[0] > 0 and nd > 1e-6
        assert np.max(np.abs(Dp - Dp.T)) < 1e-9 * nd and abs(np.trace(Dp)) < 1e-9 * nd  # plastic stretching
        S = np.linalg.inv(Dm) @ np.linalg.inv(Fp0)
        assert np.allclose(r @ S, Fe, rtol=0, atol=1e-13)
        S1 = S - np.linalg.inv(r) @ np.linalg.inv(I + Dp) @ Dp @ Fe
        ref = np.linalg.inv(Dm) @ np.linalg.inv(Fp1)
        assert np.max(np.abs(S1 - ref)) < 1e-11 * np.abs(ref).max()


# --------------------------------------------------------------------------------------------
# legacy stress laws (InteractionLinearElastic / LinearElasticDeviatoric / SecondOrderNeoHookean /
# PoroNeoHookean .cl, restated with the node-0 base)
# --------------------------------------------------------------------------------------------


def legacy_forces(law, P, X, x, tets=None):
    _, t, ev, _ = one_tet(X, tets)
    pts = O.soa(x)
    f = np.zeros_like(pts)
    vm, pr = np.zeros(1), np.zeros(1)
    O.legacy_force(law, P, pts, t, ev, f, vm, pr)
    return f.reshape(3, -1).T, vm[0], pr[0]


def hooke(F, mu, lam):
    eps = 0.5 * (F + F.T) - np.eye(3)
    return 2 * mu * eps + lam * np.trace(eps) * np.eye(3)


def test_legacy_linear_laws_closed_form_and_cross_identity():
    """The face forces carry the law's Cauchy stress exactly (virial), Hooke's law for the
    linear law, and the deviatoric law with K = lambda + 2 mu / 3 is the linear law."""
    E, nu = 0.01, 0.4
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    K = lam + 2 * mu / 3
    for _ in range(4):
        F = random_F(0.2)
        x = deform(F, REF_TET, RNG.normal(size=3))
        f0, vm0, pr0 = legacy_forces(0, O.LegacyParams(E=E, nu=nu), REF_TET, x)
        f1, vm1, pr1 = legacy_forces(1, O.LegacyParams(shear=mu, bulk=K), REF_TET, x)
        sig = hooke(F, mu, lam)
        assert np.allclose(virial_stress(REF_TET, x, f0), sig, rtol=1e-10, atol=1e-15)
        assert np.allclose(f1, f0, rtol=1e-10, atol=1e-16)
        dev = sig - np.trace(sig) / 3 * np.eye(3)
        assert vm0 == pytest.approx(np.sqrt(1.5 * np.sum(dev * dev)), rel=1e-10)
        assert pr0 == pytest.approx(np.trace(sig) / 3, rel=1e-10)
        assert np.allclose(f0.sum(0), 0, atol=1e-15)
    # small strain: the forces are the gradient of the small-strain energy up to O(strain^2)
    x = deform(np.eye(3) + 1e-4 * RNG.uniform(-1, 1, (3, 3)), REF_TET)
    f0 = legacy_forces(0, O.LegacyParams(E=E, nu=nu), REF_TET, x)[0]

    def w(F):
        eps = 0.5 * (F + F.T) - np.eye(3)
        return mu * np.sum(eps * eps) + 0.5 * lam * np.trace(eps) ** 2

    assert np.allclose(f0, energy_forces(w, REF_TET, x, h=1e-7), rtol=2e-3, atol=1e-3 * np.abs(f0).max())


def test_legacy_second_order_neo_hookean_energy_consistency():
    mu, mu2, K = 0.004, 0.003, 0.02
    P = O.LegacyParams(shear=mu, bulk=K, shear2=mu2)

    def psi(F):  # mu/2 (I1b - 3) + mu2/4 (I1b - 3)^2 + K/2 (J - 1)^2
        J = np.linalg.det(F)
        I1b = J ** (-2 / 3) * np.trace(F.T @ F)
        return 0.5 * mu * (I1b - 3) + 0.25 * mu2 * (I1b - 3) ** 2 + 0.5 * K * (J - 1) ** 2

    for _ in range(4):
        x = deform(random_F(0.25), REF_TET) + 0.03 * RNG.normal(size=(4, 3))
        f = legacy_forces(2, P, REF_TET, x)[0]
        assert np.allclose(f, energy_forces(psi, REF_TET, x), rtol=1e-6, atol=1e-9 * np.abs(f).max())
    f, vm, pr = legacy_forces(2, P, REF_TET, REF_TET.copy())
    assert np.max(np.abs(f)) < 1e-15 and vm == 0 and pr == 0


def test_legacy_poro_neo_hookean_energy_consistency():
    """sigma = a B + b I with J sigma = 2 (dPsi/dI1) B + J (dPsi/dJ) I for Psi = I1 g(J),
    g = a J / 2: the b of the law is I1 g'(J) exactly when the law is hyperelastic."""
    mu, phi = 0.004, 0.3

    def g(J):
        eta = 1 + (J - 1) / phi
        return 0.25 * mu * (4 - 2 / J - 2 * (phi + 2 * (J - 1)) / (J ** (2 / 3) * eta ** (1 / 3)))

    def psi(F):
        return np.trace(F.T @ F) * g(np.linalg.det(F))

    P = O.LegacyParams(shear=mu, phi=phi)
    for _ in range(4):
        x = deform(random_F(0.2), REF_TET) + 0.02 * RNG.normal(size=(4, 3))
        f = legacy_forces(3, P, REF_TET, x)[0]
        assert np.allclose(f, energy_forces(psi, REF_TET, x), rtol=1e-6, atol=1e-9 * np.abs(f).max())
    f, vm, pr = legacy_forces(3, P, REF_TET, REF_TET.copy())
    assert np.max(np.abs(f)) < 1e-15
    # objectivity: a rigid rotation of the deformed state rotates the forces
    Q = scipy.linalg.expm(np.array([[0, -0.3, 0.2], [0.3, 0, -0.1], [-0.2, 0.1, 0]]))
    x = deform(random_F(0.2), REF_TET)
    fa = legacy_forces(3, P, REF_TET, x)[0]
    fb = legacy_forces(3, P, REF_TET, x @ Q.T)[0]
    assert np.allclose(fb, fa @ Q.T, rtol=1e-10, atol=1e-16)


def test_linear_viscoelastic_prony_relaxation():
    """InteractionLinearViscoelastic.cl: a shear strain held from step 1 relaxes the deviatoric
    stress as the Prony series 2 mu (w_inf + sum_j w_j exp(-(k-1)/tau_j)) eps (the instantaneous
    modulus at step 1); without terms the law is the linear-elastic law."""
    E, nu = 0.01, 0.4
    mu = E / (2 * (1 + nu))
    w, tau = [0.3, 0.2], [5.0, 40.0]
    P = O.lve_params(E, nu, w, tau)
    g = 0.02
    F = np.array([[1, g, 0], [0, 1, 0], [0, 0, 1.0]])
    x = deform(F, REF_TET)
    _, t, ev, _ = one_tet(REF_TET)
    devold, trans = np.zeros(9), np.zeros(18)
    eps12 = 0.5 * g
    for k in range(1, 30):
        f = np.zeros(12)
        O.linear_viscoelastic(P, O.soa(x), t, ev, f, devold, trans)
        sig = virial_stress(REF_TET, x, f.reshape(3, -1).T)
        G = mu * (1 - sum(w) + sum(wj * np.exp(-(k - 1) / tj) for wj, tj in zip(w, tau)))
        assert sig[0, 1] == pytest.approx(2 * G * eps12, rel=1e-9), k
    # no terms: the linear-elastic law
    x = deform(random_F(0.1), REF_TET)
    f0 = legacy_forces(0, O.LegacyParams(E=E, nu=nu), REF_TET, x)[0]
    f = np.zeros(12)
    O.linear_viscoelastic(O.lve_params(E, nu, [], []), O.soa(x), t, ev, f, np.zeros(9), np.zeros(0) + np.zeros(1))
    assert np.allclose(f.reshape(3, -1).T, f0, rtol=1e-10, atol=1e-16)


# --------------------------------------------------------------------------------------------
# legacy contacts and integrators (RoundTip, AdhesivePlane, PointZeroForce, CalcCOM/COMForce,
# OverDamped, velocityVerlet2, SphereIntegral/TipIntegral, AdhesivePlaneSurfaceIntegral)
# --------------------------------------------------------------------------------------------


def test_round_tip_closed_form():
    R, ypos = 2.0, 5.0
    pts = np.array([[0.0, 1.0, 0.0], [3.0, 5.5, 0.0], [0.5, 4.0, 0.2], [0.0, 9.0, 4.0]])
    f = np.zeros(12)
    O.round_tip(f, O.soa(pts), R, ypos)
    f = f.reshape(3, -1).T
    # below the centre: sphere; |r| = 4 -> 0.1 (4 - 2) along -y
    assert np.allclose(f[0], [0, -0.2, 0])
    # above the centre: cylinder (r_y = 0), |r| = 3 -> 0.1 (3 - 2) along +x
    assert np.allclose(f[1], [0.1, 0, 0])
    # inside the sphere: max(|r| - R, 0) = 0
    assert np.allclose(f[2], 0)
    assert np.allclose(f[3], [0, 0, 0.1 * (4 - 2)])


def test_adhesive_plane_is_the_lennard_jones_wall():
    eps, rm, wall = 0.3, 0.8, -1.0
    sigma = 2 ** (-1 / 6) * rm
    d = np.array([rm, 0.7, 1.1, 2.0])
    pts = np.stack([np.zeros(4), wall + d, np.zeros(4)], 1)
    on = np.array([1, 1, 1, 0], np.int32)
    f = np.zeros(12)
    O.adhesive_plane(f, O.soa(pts), wall, eps, rm, on)
    fy = f.reshape(3, -1)[1]
    U = lambda x: 4 * eps * ((sigma / x) ** 12 - (sigma / x) ** 6)
    h = 1e-6
    assert abs(fy[0]) < 1e-12  # the minimum of the potential sits at rm
    for i in (1, 2):
        assert fy[i] == pytest.approx(-(U(d[i] + h) - U(d[i] - h)) / (2 * h), rel=1e-6)
    assert fy[3] == 0  # not a surface point


def test_point_zero_and_center_of_mass_forces():
    pts = RNG.normal(size=(7, 3))
    P = O.soa(pts)
    f = np.zeros(21)
    O.point_zero_force(f, P, 0.5)
    f = f.reshape(3, -1)
    assert np.allclose(f[0], -0.5 * pts[0, 0]) and np.allclose(f[1], 0) and np.allclose(f[2], -0.5 * pts[0, 2])
    com = O.center_of_mass(P)
    assert np.allclose(com, [pts[:, 0].mean(), pts[:, 2].mean()])
    g = np.zeros(21)
    O.com_force(g, 2.0, com)
    g = g.reshape(3, -1)
    assert np.allclose(g[0], -2 * com[0]) and np.allclose(g[1], 0) and np.allclose(g[2], -2 * com[1])


def test_over_damped_and_velocity_verlet2_recurrences():
    N = 5
    x = RNG.normal(size=3 * N)
    f = RNG.normal(size=3 * N)
    x0, f0 = x.copy(), f.copy()
    O.over_damped(x, f, point3fix=True)
    d = f0.copy()
    d[3] = d[2 * N + 3] = 0  # node 3 keeps x and z
    assert np.allclose(x, x0 + d) and np.all(f == 0)
    # velocityVerlet2, no damping, constant acceleration a from rest: x_n = x_0 + a n (n + 1) / 2
    m = np.full(N, 2.0)
    a = RNG.normal(size=3 * N)
    x, v, fo = np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N)
    for n in range(1, 8):
        F = a * np.repeat(m[None], 3, 0).ravel()
        O.velocity_verlet2(x, v, F, fo, m, 0.0)
        assert np.allclose(x, a * n * (n + 1) / 2)
    # with damping xi the velocity tends to a / xi
    x, v, fo = np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N)
    for _ in range(3000):
        O.velocity_verlet2(x, v, a * 2.0, fo, m, 0.05)
    assert np.allclose(v, a / 0.05, rtol=1e-6)


def _tet_samples(P, n, rng):
    w = rng.dirichlet(np.ones(4), size=n)
    return w @ P


def test_volume_integral_contacts():
    """Bower's 4-point rule, scaled by |det r| / 4 as the kernels write it: the nodal forces sum
    to 6 times the volume integral of the force density, their first moment to 6 times its first
    moment; the density itself at closed-form points (sphere below the centre; the tip's cone)."""
    rng = np.random.default_rng(3)
    P = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.0, 0.1], [0.1, 0.2, 0.9]]) + [0.3, -3.0, 0.2]
    R, ypos, k = 1.0, 0.5, 0.7

    def dens(x):
        r = x - [0, ypos, 0]
        L = np.linalg.norm(r, axis=1)
        return np.exp(-k * (L - R))[:, None] * r / L[:, None]

    f = np.zeros(12)
    O.volume_quadrature(0, f, O.soa(P), O.soa(np.array([[0, 1, 2, 3]]), np.uint32), R, 0.0, (0.0, ypos, 0.0), k)
    f = f.reshape(3, -1).T
    V = abs(np.linalg.det((P[:3] - P[3]).T)) / 6
    xs = _tet_samples(P, 400000, rng)
    dx = dens(xs)
    assert np.allclose(f.sum(0), 6 * V * dx.mean(0), rtol=3e-3, atol=1e-4 * np.abs(f).max())
    mom = sum(np.outer(f[a], P[a]) for a in range(4))
    mom_ref = 6 * V * (dx[:, :, None] * xs[:, None, :]).mean(0)
    # first moment: sum_a F_a x_a = (6 V) E[f(x) x] for the P1 interpolation weights (to quadrature accuracy)
    assert np.allclose(mom, mom_ref, rtol=2e-2, atol=2e-2 * np.abs(mom_ref).max())
    # a tiny tetra: the density at its centre, closed forms
    small = 1e-3 * np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    V6 = abs(np.linalg.det((small[:3] - small[3]).T))
    alpha = 0.3
    tip = np.array([0.0, ypos, 0.0])
    for tip_shape, c in [(0, np.array([0.0, -2.0, 0.0])), (1, np.array([0.0, -3.0, 0.0])),
                         (1, np.array([4.0, 2.0, 0.0]))]:
        Q = small + c
        cc = Q.mean(0)
        if tip_shape == 1 and cc[1] >= ypos - np.sin(alpha) * np.hypot(cc[0], cc[2]):
            # beside the tip, above the cone line: n = cos a e_rho - sin a e_y, dis = x.n - tip.n - R
            e = np.array([cc[0], 0.0, cc[2]]) / np.hypot(cc[0], cc[2])
            n = np.cos(alpha) * e + np.array([0.0, -np.sin(alpha), 0.0])
            expect = np.exp(-k * (cc @ n - tip @ n - R)) * n
        else:  # sphere: below the centre of the sphere (or of the tip's rounded end)
            r = cc - tip
            expect = np.exp(-k * (np.linalg.norm(r) - R)) * r / np.linalg.norm(r)
        g = np.zeros(12)
        O.volume_quadrature(tip_shape, g, O.soa(Q), O.soa(np.array([[0, 1, 2, 3]]), np.uint32), R, alpha,
                            tuple(tip), k)
        total = g.reshape(3, -1).T.sum(0) / (1.5 * V6 / 6 * 4)  # 0.25 |det| sum_q f_q = 1.5 V6/6 * 4 * mean f
        assert np.allclose(total, expect, rtol=2e-3, atol=1e-6), (tip_shape, total, expect)


def test_adhesive_plane_surface_integral():
    eps, rm, wall = 0.2, 0.8, -1.0
    sigma = 2 ** (-1 / 6) * rm
    lj = lambda d: 24 * eps / d * (2 * (sigma / d) ** 12 - (sigma / d) ** 6)
    # a triangle parallel to the wall: constant density, area / 3 to each node
    P = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]) + [0, wall + 0.9, 0]
    f = np.zeros(9)
    O.adhesive_plane_surface(f, O.soa(P), O.soa(np.array([[0, 1, 2]]), np.uint32), wall, eps, rm)
    fy = f.reshape(3, -1)[1]
    assert np.allclose(fy, 0.5 / 3 * lj(0.9), rtol=1e-12)
    # tilted: the sum is the area integral of the density (the 3-point rule is exact to degree 2;
    # the density varies slowly here)
    Q = np.array([[0.0, 2.0, 0.0], [1.0, 2.3, 0.0], [0.0, 2.1, 1.0]])
    g = np.zeros(9)
    O.adhesive_plane_surface(g, O.soa(Q), O.soa(np.array([[0, 1, 2]]), np.uint32), wall, eps, rm)
    rng = np.random.default_rng(4)
    w = rng.dirichlet(np.ones(3), size=400000)
    ys = w @ Q[:, 1]
    area = 0.5 * np.linalg.norm(np.cross(Q[0] - Q[2], Q[1] - Q[2]))
    assert g.reshape(3, -1)[1].sum() == pytest.approx(area * lj(ys - wall).mean(), rel=2e-3)


def test_hard_adhesive_plane_surface_integral():
    """exp(-k d) - A exp(-d) by the 3-point rule: a triangle parallel to the wall gets area/3 of the
    density per node; the total force is -dE/dh of the energy area * U(d), U = exp(-k d)/k - A exp(-d),
    written independently; the force vanishes at d* = ln(A)/(1 - k); a tilted triangle integrates the
    density (Monte Carlo)."""
    A, k, wall = 0.6, 3.0, -1.0
    dens = lambda d: np.exp(-k * d) - A * np.exp(-d)  # noqa: E731
    U = lambda d: np.exp(-k * d) / k - A * np.exp(-d)  # noqa: E731
    tri = O.soa(np.array([[0, 1, 2]]), np.uint32)
    base = np.array([[0.0, 0.0, 0.0], [1.3, 0.0, 0.0], [0.0, 0.0, 0.7]])
    area = 0.5 * 1.3 * 0.7

    def fy(P):
        f = np.zeros(9)
        O.hard_adhesive_plane_surface(f, O.soa(P), tri, wall, A, k)
        return f.reshape(3, -1)[1]

    for h in (0.2, 0.9, 2.5):
        P = base + [0, wall + h, 0]
        assert np.allclose(fy(P), area / 3 * dens(h), rtol=1e-12)
        e = 1e-5
        dE = (area * U(h + e) - area * U(h - e)) / (2 * e)
        assert fy(P).sum() == pytest.approx(-dE, rel=1e-8)
    dstar = np.log(A) / (1 - k)
    assert np.abs(fy(base + [0, wall + dstar, 0])).max() < 1e-14
    Q = np.array([[0.0, 0.6, 0.0], [1.0, 0.9, 0.0], [0.0, 0.7, 1.0]])
    g = fy(Q)
    rng = np.random.default_rng(6)
    w = rng.dirichlet(np.ones(3), size=400000)
    area_q = 0.5 * np.linalg.norm(np.cross(Q[0] - Q[2], Q[1] - Q[2]))
    assert g.sum() == pytest.approx(area_q * dens(w @ Q[:, 1] - wall).mean(), rel=3e-3)


# --------------------------------------------------------------------------------------------
# contacts, integrator, tip force
# --------------------------------------------------------------------------------------------


def test_contacts_closed_form():
    k, R, yc = 0.1, 2.0, 5.0
    x = np.array([[0.0, 4.0, 0.0], [0.3, 10.0, 0.0], [1.0, 5.0, 1.0], [0.0, -3.0, 0.0]])
    pts = O.soa(x)
    f = np.zeros_like(pts)
    O.sphere(k, f, pts, R, yc)
    f = f.reshape(3, -1).T
    assert np.allclose(f[0], [0, -k * 1.0, 0])  # 1 deep, pushed away from the tip centre (down)
    assert np.allclose(f[1], 0)  # outside
    d = np.sqrt(2.0)
    assert np.allclose(f[2], k * (R - d) * np.array([1, 0, 1]) / d)
    f = np.zeros_like(pts)
    O.substrate(k, f, pts, -2.0)
    assert np.allclose(f.reshape(3, -1).T[:, 1], [0, 0, 0, k * 1.0])
    f = np.zeros_like(pts)
    O.plane(k, f, pts, 4.5)
    assert np.allclose(f.reshape(3, -1).T[:, 1], [0, -k * 5.5, -k * 0.5, 0])


def test_velocity_verlet_recurrence_and_terminal_velocity():
    xi = 0.005
    a, b = 1 - xi / 2, 1 + xi / 2
    rho = a / b
    # spring f = -k x on one node: x_{n+1} = (1+rho) x_n - rho x_{n-1} + f_n / b for n >= 1
    kk = 0.3
    x = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.0])  # N = 2, SoA
    v, f, fo = np.zeros(6), np.zeros(6), np.zeros(6)
    xs = []
    for _ in range(50):
        f[:] = -kk * x
        xs.append(x[0])
        O.velocity_verlet(x, v, f, fo)
    xs = np.array(xs)
    for n in range(1, 48):
        assert xs[n + 1] == pytest.approx((1 + rho) * xs[n] - rho * xs[n - 1] - kk * xs[n] / b, abs=1e-14)
    # constant force: velocity tends to f / xi
    x, v, fo = np.zeros(6), np.zeros(6), np.zeros(6)
    for _ in range(20000):
        f = np.full(6, 1e-4)
        O.velocity_verlet(x, v, f, fo)
        assert np.all(f == 0)  # force buffer cleared (VelocityVerlet.cl:246-248)
    assert v[0] == pytest.approx(1e-4 / xi, rel=1e-6)


def test_fix_top_bottom():
    N = 6
    x = np.arange(3 * N, dtype=np.float64) + 1
    v, f, fo = np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N)
    O.velocity_verlet(x, v, f, fo, fix_top_bottom=True)
    X = x.reshape(3, N).T
    assert X[3, 0] == 0 and X[3, 2] == 0 and X[4, 0] == 0 and X[4, 2] == 0 and X[3, 1] != 0 and X[2, 0] != 0


def test_tip_force_brute_force():
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, _ = generate_sphere_mesh(4.0, 0.8, seed=3)
    R, k, yc = 1.5, 0.1, 4.5
    fz, mx = O.tip_force(O.soa(P), 1, k, R, yc)
    ref = O.force_si_sphere(P, yc, R, k, 1.0, 1.0)
    assert fz == pytest.approx(ref, rel=1e-12)
    assert fz > 0  # cell pushes the tip up
    d = np.maximum(R - np.linalg.norm(P - [0, yc, 0], axis=1), 0).max()
    assert mx == pytest.approx(d)


def test_run_momentum_decay_without_contacts():
    """no external force: total momentum obeys sum v_{n+1} = rho * sum v_n."""
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, T = generate_sphere_mesh(3.0, 0.75, seed=11)
    prm = O.RunParams()
    prm.model, prm.E, prm.nu, prm.ratio, prm.nthreads = O.MR, 0.01, 0.45, 1.0, 1
    st = O.OracleState(P * 2.0, T, prm)  # reference from unscaled P: prestressed start
    st.ev[:] = O.soa(O.edge_vectors_volumes(P * 1.9, T)[0])
    st.vol[:] = O.edge_vectors_volumes(P * 1.9, T)[1]
    st.vel[:] = 1e-3
    p_prev = st.vel.reshape(3, -1).sum(1)
    st.run(1, record=False)
    rho = (1 - 0.0025) / (1 + 0.0025)
    p1 = st.vel.reshape(3, -1).sum(1)
    assert np.allclose(p1, rho * p_prev, rtol=1e-9)


def test_explicit_stability_bound_of_the_cell_units():
    """The batched cells' unit choice keeps the largest stiffness eigenvalue (unit node mass) below
    the explicit limit 4: power iteration on the oracle's force Jacobian for the stiffest sampled
    material (nu = 0.49) at the chosen lattice size."""
    from nofluidx3d_amd.batch import CellSpec, sim_points
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    s = CellSpec(2.5, 0.5, 3002, 3000.0, 0.49, None, 1.5, 20.0)
    P, T = generate_sphere_mesh(s.radius_um, s.spacing_um, s.seed)
    x = sim_points(P, s.units(800)["RadiusSIM"])
    ev, vol = O.edge_vectors_volumes(x, T)
    args = (O.soa(T, np.uint32), O.soa(ev), np.ascontiguousarray(vol))

    def forces(X):
        f = np.zeros(3 * len(X))
        O.tetra_force(O.MR, 0.01, s.nu, 1.0, O.soa(X), *args, f)
        return f.reshape(3, -1).T

    v = np.random.default_rng(0).normal(size=x.shape)
    lam = 0.0
    for _ in range(60):
        v /= np.linalg.norm(v)
        Kv = -(forces(x + 1e-6 * v) - forces(x - 1e-6 * v)) / 2e-6
        lam = float((v * Kv).sum())
        v = Kv
    assert lam < 2.5, lam
