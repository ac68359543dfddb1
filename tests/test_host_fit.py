"""Host logic of the surrogate-assisted fitter and the Nelder-Mead baseline, on the CPU, with an
analytic stand-in for the batched simulator and for the network (both obey the simulator's exact
E-viscosity symmetry: the normalised curve depends on viscosity / E, not on E)."""

import numpy as np

from nofluidx3d_amd.batch import CellSpec
from nofluidx3d_amd.surrogate import fit
from nofluidx3d_amd.surrogate.train import Normaliser

NC = 256
S = np.linspace(0.0, 1.0, NC)


def shape(nu, log_ratio, tip, speed, R):
    """normalised curve of an analytic 'cell' (smooth in every input)"""
    k = 0.02 * (1.0 + nu) * np.sqrt(tip / (R + tip))
    relax = 1.0 / (1.0 + 10 ** (log_ratio + 2.5) * speed / R)
    return k[:, None] * S[None, :] ** 1.5 * (1.0 - 0.6 * relax[:, None] * S[None, :])


class FakeBatch:
    steps = 10000

    def reset(self, specs):
        self.specs = specs

    def advance(self, n):
        pass

    def curves_normalized(self):
        sp = self.specs
        return shape(np.array([s.nu for s in sp]), np.log10([s.viscosity_si / s.E_si for s in sp]),
                     np.array([s.tip_radius_um for s in sp]), np.array([s.speed_um_s for s in sp]),
                     np.array([s.radius_um for s in sp]))


class FakeNet:
    def __init__(self, norm):
        self.norm = norm

    def predict(self, Xn):
        X = np.asarray(Xn) * self.norm.x_std + self.norm.x_mean
        # surrogate_inputs layout: log10 viscosity/E, nu, log10 tip, log10 speed, log10 R, log10 spacing
        return shape(X[:, 1], X[:, 0], 10 ** X[:, 2], 10 ** X[:, 3], 10 ** X[:, 4]) / self.norm.y_scale


def cells(n, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(CellSpec(rng.uniform(4, 10), 0.5, i, 10 ** rng.uniform(2.3, 3.7), rng.uniform(0.32, 0.47),
                            10 ** rng.uniform(0.8, 2.8), rng.uniform(1, 5), 10 ** rng.uniform(0, 1.5)))
    return out


def blind(cs):
    import dataclasses

    return [dataclasses.replace(c, E_si=1000.0, nu=0.4, viscosity_si=10.0) for c in cs]


def test_candidate_scores_equal_the_direct_amplitude_and_error():
    """fit._scores (dot-product form) against the direct computation (scaled curves, closed-form
    amplitude, explicit residual norm), with and without the additive correction; candidate 0 of
    each target is a near-fit (amplitude 3, residual 1 %) so the unclipped branch and small errors
    are covered."""
    rng = np.random.default_rng(11)
    n, m = 5, 7
    pred = rng.normal(size=(n, m, NC)).astype(np.float32)
    scale = 10 ** rng.uniform(-12, -9, size=(n, m))
    corr = 0.3 * rng.normal(size=(n, NC))
    for c in (None, corr):
        P = pred.astype(np.float64) + (0 if c is None else c[:, None, :])
        t = 3.0 * scale[:, :1] * P[:, 0] * (1.0 + 0.01 * rng.normal(size=(n, NC)))
        F = scale[..., None] * P
        a_ref = np.clip(np.sum(F * t[:, None], -1) / np.sum(F * F, -1), 1e-6, 1e6)
        e_ref = np.linalg.norm(a_ref[..., None] * F - t[:, None], axis=-1) / np.linalg.norm(t, axis=-1)[:, None]
        a, e = fit._scores(pred, t, scale, c)
        assert np.all(np.abs(a[:, 0] - 3.0) < 0.01) and np.all(e[:, 0] < 0.02)
        assert np.allclose(a, a_ref, rtol=1e-12, atol=0)
        assert np.allclose(e, e_ref, rtol=1e-9, atol=0)


def test_symmetry_helpers_keep_the_box_and_the_ratio():
    rng = np.random.default_rng(0)
    th = np.stack([rng.uniform(1, 5, 500), rng.uniform(0.2, 0.6, 500), rng.uniform(-1, 4, 500)], 1)
    c = fit.clip_theta(th)
    assert np.all(c >= fit.LO) and np.all(c <= fit.HI)
    inside = np.all((th >= fit.LO) & (th <= fit.HI), axis=1)
    assert np.array_equal(c[inside], th[inside])
    # the network's inputs see E and the viscosity through their ratio only (the exact symmetry)
    cs = cells(6)
    th = np.array([[np.log10(c.E_si), c.nu, np.log10(c.viscosity_si)] for c in cs])
    shifted = th + np.array([0.37, 0.0, 0.37])
    assert np.allclose(fit.features(th, cs), fit.features(shifted, cs), rtol=0, atol=1e-12)
    X = fit.features(th, cs)
    assert np.allclose(X[:, 0], th[:, 2] - th[:, 0]) and np.allclose(X[:, 5], np.log10([c.spacing_um for c in cs]))


def test_stabilize_enforces_the_relaxation_limit():
    cs = cells(20)
    T0 = fit.time_units(cs, 10000)
    th = np.array([[3.5, 0.35, 0.0]] * 20)
    st = fit.stabilize(th, T0)
    for c, t in zip(cs, st):
        spec = fit.theta_to_spec(t, c)
        assert spec.relaxation_per_step(10000) <= 0.1 * (1 + 1e-9)
    loose = np.array([[2.0, 0.35, 3.0]] * 20)
    assert np.array_equal(fit.stabilize(loose, T0), loose)


def test_surrogate_fit_recovers_parameters_with_an_exact_network():
    cs = cells(24)
    bs = FakeBatch()
    targets = fit.simulate(bs, cs)
    th = np.array([[np.log10(c.E_si), c.nu, np.log10(c.viscosity_si)] for c in cs])
    X = fit.features(th, cs)
    norm = Normaliser(X.mean(0), np.where(X.std(0) > 1e-6, X.std(0), 1.0), 0.01)
    r = fit.fit_surrogate(FakeNet(norm), norm, targets, blind(cs), bs, rounds=8, ncand=64, ntop=2)
    assert np.median(r["sim_err"]) < 2e-3, np.median(r["sim_err"])
    assert r["simulations_per_target"] == 3
    truth = np.array([[np.log10(c.E_si), c.nu, np.log10(c.viscosity_si)] for c in cs])
    # the amplitude fixes E exactly once the shape (nu, viscosity / E) is right
    assert np.median(np.abs(r["theta"][:, 0] - truth[:, 0])) < 0.02
    assert np.median(np.abs((r["theta"][:, 2] - r["theta"][:, 0]) - (truth[:, 2] - truth[:, 0]))) < 0.05


def test_nelder_mead_history_is_monotone_and_improves():
    cs = cells(16, seed=9)
    bs = FakeBatch()
    targets = fit.simulate(bs, cs)
    r = fit.fit_nelder_mead(targets, blind(cs), bs, max_evals=40)
    h = r["history"]
    assert h.shape == (r["evaluations"], 16)
    assert np.all(np.diff(h, axis=0) <= 1e-15)
    assert np.median(h[-1]) < 0.5 * np.median(h[3])
    assert np.allclose(np.min(h, axis=0), r["sim_err"])


class BiasedNet(FakeNet):
    """a network with a smooth cell-dependent error of 8 % (as the mesh, which the network does not
    see, makes the simulator's curve differ from the network's)"""

    def predict(self, Xn):
        X = np.asarray(Xn) * self.norm.x_std + self.norm.x_mean
        base = shape(X[:, 1], X[:, 0], 10 ** X[:, 2], 10 ** X[:, 3], 10 ** X[:, 4])
        return base * (1.0 + 0.08 * np.sin(6 * S[None, :] + 7.0 * X[:, 4:5])) / self.norm.y_scale


def test_corrected_round_beats_a_second_candidate_at_equal_simulations():
    """Two simulations per target either way: the two best network candidates, or the best one
    and one round corrected by its simulation residual.  The correction removes the network's
    bias, so its fits are several times closer."""
    cs = cells(32, seed=5)
    bs = FakeBatch()
    targets = fit.simulate(bs, cs)
    th = np.array([[np.log10(c.E_si), c.nu, np.log10(c.viscosity_si)] for c in cs])
    X = fit.features(th, cs)
    norm = Normaliser(X.mean(0), np.where(X.std(0) > 1e-6, X.std(0), 1.0), 0.01)
    two = fit.fit_surrogate(BiasedNet(norm), norm, targets, blind(cs), bs, ntop=2, confirm=False)
    cor = fit.fit_surrogate(BiasedNet(norm), norm, targets, blind(cs), bs, ntop=1, corrections=1, confirm=False)
    assert two["simulations_per_target"] == cor["simulations_per_target"] == 2
    assert np.median(cor["sim_err"]) < 0.3 * np.median(two["sim_err"]), (np.median(cor["sim_err"]),
                                                                         np.median(two["sim_err"]))
    # the fit's curve is the simulator's at the fitted parameters (no confirming run needed)
    again = fit.simulate(bs, [fit.theta_to_spec(t, c) for t, c in zip(cor["theta"], blind(cs))])
    assert np.max(fit.rel_err(again, cor["curve"])) < 1e-9
