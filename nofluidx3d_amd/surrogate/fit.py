"""Surrogate-assisted fitting of (E, nu, viscosity) to force-indentation curves.

A target is a measured curve F(delta) (FP64, newtons, 256 samples) of a cell whose geometry and
protocol are known (mesh, cell radius, tip radius, speed, depth).  The surrogate predicts
F / (E R^2), so a candidate (E, nu, eta) predicts F = E R^2 f(inputs).

The simulator has an exact symmetry: (E, eta) -> (a E, a eta) leaves every simulation-unit
quantity of the cell unchanged, so the simulated curve becomes exactly a F.  Hence for any curve
shape the best amplitude is a closed form, a = <F, F_t> / <F, F>, and the search runs over the
shape parameters (nu, log10 eta/E) only.

Surrogate-assisted fit (``fit_surrogate``), for all targets at once:
1. ``rounds`` cross-entropy rounds of ``ncand`` candidates (nu, log10 eta/E) per target, each
   round one batched inference over targets x ncand rows; each candidate's E comes from the
   closed-form amplitude against the target (the network is evaluated at the candidate's symmetric
   image inside the sampled E and viscosity ranges);
2. the ``ntop`` best distinct candidates per target are simulated with the batched simulator on
   the targets' own meshes and protocols; each simulated curve gets its closed-form amplitude and
   the best becomes the fit;
3. ``corrections`` corrected rounds: the residual of the best simulation against the network at
   the same parameters is added to every prediction, the search runs again around the best point
   and its best candidate is simulated;
4. optionally the fit's parameters are simulated once more (``confirm``); otherwise the fit's curve
   is the refining simulation's scaled by its amplitude, which the symmetry makes exact.
Baseline (``fit_nelder_mead``): simulation-only Nelder-Mead on (log10 E, nu, log10 eta), all
targets in lock-step, one batched simulation per evaluation.
"""

from __future__ import annotations

import dataclasses
import time

import numpy as np

from nofluidx3d_amd.batch import surrogate_inputs

LO = np.array([2.0, 0.3, 0.0])  # log10 E [Pa], nu, log10 viscosity [Pa s]
HI = np.array([4.0, 0.49, 3.0])


def theta_to_spec(theta, base):
    logE, nu, logeta = theta
    return dataclasses.replace(base, E_si=10 ** float(logE), nu=float(np.clip(nu, 0.3, 0.49)),
                               viscosity_si=10 ** float(logeta))


def features(theta, bases):
    """surrogate inputs for parameter rows theta (n, 3) of cells ``bases`` (n)"""
    return surrogate_inputs(10 ** theta[:, 0], theta[:, 1], 10 ** theta[:, 2], [b.tip_radius_um for b in bases],
                            [b.speed_um_s for b in bases], [b.radius_um for b in bases],
                            [b.spacing_um for b in bases])


def curve_force(curve_norm, spec):
    """normalised curve -> newtons"""
    return curve_norm * spec.E_si * (spec.radius_um * 1e-6) ** 2


def simulate(bs, specs):
    """Curves (n, ncurve) in newtons from the batched simulator ``bs`` (same meshes as ``specs``)."""
    bs.reset(specs)
    bs.advance(bs.steps)
    cn = bs.curves_normalized()
    return np.stack([curve_force(cn[i], s) for i, s in enumerate(specs)])


def rel_err(F, targets):
    return np.linalg.norm(F - targets, axis=-1) / np.linalg.norm(targets, axis=-1)


def amplitude(F, targets):
    """closed-form best scale a of curves F (..., n) against targets (..., n)"""
    return np.sum(F * targets, axis=-1) / np.maximum(np.sum(F * F, axis=-1), 1e-300)


def clip_theta(theta):
    """Keep log10 E and log10 viscosity inside the box by a common shift (the symmetry direction),
    then clip every coordinate."""
    th = np.array(theta, np.float64, copy=True)
    lo = np.maximum(LO[0] - th[..., 0], LO[2] - th[..., 2])
    hi = np.minimum(HI[0] - th[..., 0], HI[2] - th[..., 2])
    shift = np.clip(0.0, lo, np.maximum(lo, hi))
    th[..., 0] += shift
    th[..., 2] += shift
    return np.clip(th, LO, HI)


def time_units(bases, steps):
    """T0 [s per step] of each cell (independent of its material)"""
    return np.array([b.units(steps)["T0"] for b in bases])


def stabilize(theta, T0, max_relax=0.1):
    """Raise the viscosity where the explicit Maxwell update would remove more than ``max_relax``
    of the deviatoric stress per step (G T0 / eta > max_relax: unstable flow update), the limit
    the training data were drawn under.  ``T0``: time unit of each row's cell."""
    th = np.array(theta, np.float64, copy=True)
    lim = th[..., 0] + np.log10(np.asarray(T0) / (2.0 * (1.0 + th[..., 1]) * max_relax))
    th[..., 2] = np.maximum(th[..., 2], lim)
    return th


def _scores(pred, targets, scale, corr=None):
    """Closed-form amplitude a (clipped to [1e-6, 1e6]) and relative L2 error of a F against the
    targets for candidate curves F = scale (pred + corr): pred (n, m, k) float32, targets (n, k),
    scale (n, m), corr (n, k) or None.  Both follow from the dot products p.t and p.p (and t.t):
    ||a F - t||^2 = a^2 F.F - 2 a F.t + t.t, so no (n, m, k) temporaries are formed."""
    t = np.asarray(targets, np.float64)
    p64 = pred.astype(np.float64)
    pt = np.matmul(p64, t[:, :, None])[..., 0]
    pp = np.einsum("ijk,ijk->ij", p64, p64)
    if corr is not None:
        c = np.asarray(corr, np.float64)
        pt = pt + np.einsum("ik,ik->i", c, t)[:, None]
        pp = pp + 2.0 * np.matmul(p64, c[:, :, None])[..., 0] + np.einsum("ik,ik->i", c, c)[:, None]
    tt = np.einsum("ik,ik->i", t, t)[:, None]
    ft, ff = scale * pt, scale * scale * pp
    a = np.clip(ft / np.maximum(ff, 1e-300), 1e-6, 1e6)
    err = np.sqrt(np.maximum(a * a * ff - 2.0 * a * ft + tt, 0.0) / tt)
    return a, err


def _search(net, norm, targets, bases, R2, T0r, rounds, ncand, rng, corr=None, start=None):
    """Cross-entropy search over the shape parameters (nu, log10 eta/E) of every target at once.
    ``corr`` (n, ncurve): additive correction of the normalised prediction per target; ``start``
    (mu, sd), each (n, 2): the first round's sampling distribution (else uniform over the box).
    Returns every candidate (n, rounds*ncand, 3) with its amplitude-fitted E, their errors against
    the targets and the inference time."""
    n = len(targets)
    slo, shi = np.array([LO[1], LO[2] - HI[0]]), np.array([HI[1], HI[2] - LO[0]])  # nu, log10 eta/E
    rep = [b for b in bases for _ in range(ncand)]
    all_theta = np.zeros((n, 0, 3))
    all_err = np.zeros((n, 0))
    infer_s = 0.0
    for r in range(rounds):
        if r == 0 and start is None:
            sh = slo + (shi - slo) * rng.random((n, ncand, 2))
        else:
            if r == 0:
                mu, sd = start
            else:
                order = np.argsort(all_err, axis=1)[:, :8]
                elite = np.take_along_axis(all_theta, order[:, :, None], 1)
                es = np.stack([elite[..., 1], elite[..., 2] - elite[..., 0]], -1)
                mu, sd = es.mean(1), es.std(1) + 0.02 * (shi - slo)
            sh = np.clip(mu[:, None, :] + sd[:, None, :] * rng.normal(size=(n, ncand, 2)), slo, shi)
        # E from the closed-form amplitude of the predicted shape.  The network sees the symmetric
        # image of each candidate inside the sampled box (same nu and viscosity / E); the candidate
        # itself keeps the amplitude's E, which may leave the box (the simulator takes any E).
        th = np.zeros((n, ncand, 3))
        th[..., 1] = sh[..., 0]
        th[..., 0] = 3.0
        th[..., 2] = th[..., 0] + sh[..., 1]
        th = stabilize(th.reshape(-1, 3), T0r).reshape(n, ncand, 3)
        for _ in range(2):
            img = clip_theta(th)
            ti = time.perf_counter()
            pred = net.predict(norm.x(features(img.reshape(-1, 3), rep))) * np.float32(norm.y_scale)
            infer_s += time.perf_counter() - ti
            a, err = _scores(pred.reshape(n, ncand, -1), targets, 10 ** img[..., 0] * R2[:, None], corr)
            th = img.copy()
            th[..., 0] += np.log10(a)
            th[..., 2] += np.log10(a)
        all_theta = np.concatenate([all_theta, th], 1)
        all_err = np.concatenate([all_err, err], 1)
    return all_theta, all_err, infer_s


def _distinct_best(all_theta, all_err, k):
    """the k best candidates per target that differ by more than 1e-3 of the box in some
    coordinate (the last one repeated where fewer exist)"""
    n = len(all_theta)
    order = np.argsort(all_err, axis=1)
    top = np.zeros((n, k, 3))
    for i in range(n):
        picked = []
        for j in order[i]:
            cand = all_theta[i, j]
            if all(np.max(np.abs(cand - p) / (HI - LO)) > 1e-3 for p in picked):
                picked.append(cand)
            if len(picked) == k:
                break
        while len(picked) < k:
            picked.append(picked[-1])
        top[i] = np.array(picked)
    return top


def fit_surrogate(net, norm, targets, bases, bs, rounds=8, ncand=64, ntop=4, corrections=0, corr_rounds=4, seed=0,
                  log=None, confirm=True):
    """targets (n, 256) newtons; bases (n) CellSpec with the known geometry/protocol; ``bs`` a
    BatchSimulation or CellGroups over the targets' meshes.  Returns dict with the fitted parameters, the fit's
    simulated curve and its error, and timings.

    ``corrections``: after the ``ntop`` refining simulations, that many corrected rounds.  Each
    adds the residual of the best simulation so far against the network at the same parameters
    (normalised curves; it carries what the network cannot see, chiefly the target's own mesh) to
    every prediction, searches again (``corr_rounds`` cross-entropy rounds around the best point)
    and simulates the new best candidate: a zeroth-order additive correction of the surrogate by
    the simulator, as in multi-fidelity model management.

    ``confirm``: simulate the final parameters once more; without it the curve is the refining
    simulation's, exactly scaled by the closed-form amplitude (identical to the simulator's output
    at the final parameters by the E-viscosity symmetry)."""
    n = len(targets)
    rng = np.random.Generator(np.random.PCG64(seed))
    R2 = np.array([(b.radius_um * 1e-6) ** 2 for b in bases])
    T0 = time_units(bases, bs.steps)
    T0r = np.repeat(T0, ncand)
    t0 = time.perf_counter()
    all_theta, all_err, infer_s = _search(net, norm, targets, bases, R2, T0r, rounds, ncand, rng)
    t_sur = time.perf_counter() - t0
    top = _distinct_best(all_theta, all_err, ntop)
    t1 = time.perf_counter()
    best_err = np.full(n, np.inf)
    theta = top[:, 0].copy()
    curve = np.zeros_like(targets)
    best_norm = np.zeros_like(targets)  # normalised simulated curve at the best parameters

    def refine(cands, label):
        nonlocal theta, curve, best_err, best_norm
        cands = stabilize(cands, T0)
        F = simulate(bs, [theta_to_spec(cands[i], bases[i]) for i in range(n)])
        cn = bs.curves_normalized()
        a = np.clip(amplitude(F, targets), 1e-6, 1e6)
        scaled = cands + np.log10(a)[:, None] * np.array([1.0, 0.0, 1.0])
        cand = stabilize(scaled, T0)
        # the amplitude-corrected parameters give exactly a F (the E-viscosity symmetry; the
        # stability limit is invariant under it); a candidate that the limit did move keeps its
        # simulated curve unscaled
        exact = np.all(np.abs(cand - scaled) <= 1e-12, axis=1)
        Fk = np.where(exact[:, None], F * a[:, None], F)
        cand = np.where(exact[:, None], cand, cands)
        e = rel_err(Fk, targets)
        better = e < best_err
        best_err[better] = e[better]
        theta[better] = cand[better]
        curve[better] = Fk[better]
        best_norm[better] = cn[better]
        if log:
            log(f"refine {label}: {time.perf_counter() - t1:.1f}s, median error {np.median(best_err):.4f}")

    for k in range(ntop):
        refine(top[:, k], f"{k + 1}/{ntop}")
    for c in range(corrections):
        rep = list(bases)
        img = clip_theta(theta)
        ti = time.perf_counter()
        at = net.predict(norm.x(features(img, rep))) * norm.y_scale
        infer_s += time.perf_counter() - ti
        corr = best_norm - at
        mu = np.stack([theta[:, 1], theta[:, 2] - theta[:, 0]], 1)
        sd = np.repeat(np.array([[0.02, 0.15]]), n, 0)
        ts = time.perf_counter()
        th_c, err_c, inf_c = _search(net, norm, targets, bases, R2, T0r, corr_rounds, ncand, rng, corr=corr,
                                     start=(mu, sd))
        infer_s += inf_c
        t_sur += time.perf_counter() - ts
        refine(_distinct_best(th_c, err_c, 1)[:, 0], f"correction {c + 1}/{corrections}")
    t_ref = time.perf_counter() - t1 - (t_sur - (t1 - t0))
    t2 = time.perf_counter()
    if confirm:
        curve = simulate(bs, [theta_to_spec(theta[i], bases[i]) for i in range(n)])
    t_conf = time.perf_counter() - t2
    return {"theta": theta, "curve": curve, "sim_err": rel_err(curve, targets),
            "predicted_refined_err": best_err, "surrogate_s": t_sur, "inference_s": infer_s,
            "refine_s": t_ref, "confirm_s": t_conf, "total_s": time.perf_counter() - t0,
            "candidates_per_target": (rounds + corrections * corr_rounds) * ncand,
            "simulations_per_target": ntop + corrections + (1 if confirm else 0)}


def fit_nelder_mead(targets, bases, bs, max_evals=60, log=None):
    """Simulation-only Nelder-Mead (coefficients 1, 2, 0.5, 0.5) on (log10 E, nu, log10 eta) in
    box-normalised coordinates, all targets in lock-step: every evaluation is one batched
    simulation of all targets.  ``history`` holds the best error of every target after each
    evaluation."""
    n = len(targets)
    T0 = time_units(bases, bs.steps)

    def evaluate(u):  # u (n, 3) in [0, 1]^3
        th = stabilize(LO + (HI - LO) * np.clip(u, 0, 1), T0)
        F = simulate(bs, [theta_to_spec(th[i], bases[i]) for i in range(n)])
        return rel_err(F, targets)

    t0 = time.perf_counter()
    hist = []

    def record(f):
        hist.append(np.min(f, axis=1) if f.ndim == 2 else f)

    simplex = np.array([[0.5, 0.5, 0.5], [0.8, 0.5, 0.5], [0.5, 0.8, 0.5], [0.5, 0.5, 0.8]])
    S = np.repeat(simplex[None], n, 0)  # (n, 4, 3)
    cols = []
    for j in range(4):
        cols.append(evaluate(S[:, j]))
        record(np.stack(cols, 1))
    f = np.stack(cols, 1)
    evals = 4
    while evals + 2 <= max_evals:
        o = np.argsort(f, 1)
        S = np.take_along_axis(S, o[:, :, None], 1)
        f = np.take_along_axis(f, o, 1)
        c = S[:, :3].mean(1)
        xr = c + (c - S[:, 3])
        fr = evaluate(xr)
        record(np.minimum(f.min(1), fr))
        expand = fr < f[:, 0]
        # one more evaluation per target: the expansion point where it may pay, else the contraction
        x2 = np.where(expand[:, None], c + 2 * (c - S[:, 3]), c + 0.5 * (S[:, 3] - c))
        f2 = evaluate(x2)
        record(np.minimum(np.minimum(f.min(1), fr), f2))
        evals += 2
        shrink = np.zeros(n, bool)
        for i in range(n):
            if expand[i]:
                S[i, 3], f[i, 3] = (x2[i], f2[i]) if f2[i] < fr[i] else (xr[i], fr[i])
                continue
            # both points were evaluated: keep the better of the acceptable reflection
            # (better than the second worst) and the acceptable contraction (better than the worst)
            ok_r, ok_c = fr[i] < f[i, 2], f2[i] < f[i, 3]
            if ok_c and (not ok_r or f2[i] < fr[i]):
                S[i, 3], f[i, 3] = x2[i], f2[i]
            elif ok_r:
                S[i, 3], f[i, 3] = xr[i], fr[i]
            else:
                shrink[i] = True
        if shrink.any() and evals + 3 <= max_evals:  # shrink towards the best vertex, re-evaluated
            S[shrink, 1:] = S[shrink, :1] + 0.5 * (S[shrink, 1:] - S[shrink, :1])
            for j in range(1, 4):
                fj = evaluate(S[:, j])
                f[shrink, j] = fj[shrink]
                record(f.min(1))
            evals += 3
        if log:
            log(f"nelder-mead evals {evals}: best median err {np.median(f.min(1)):.4f}")
    j = np.argmin(f, 1)
    theta = stabilize(LO + (HI - LO) * np.clip(S[np.arange(n), j], 0, 1), T0)
    return {"theta": theta, "sim_err": f[np.arange(n), j], "total_s": time.perf_counter() - t0, "evaluations": evals,
            "history": np.array(hist)}
