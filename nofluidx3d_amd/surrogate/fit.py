"""Surrogate-assisted fitting of (E, nu, viscosity) to force-indentation curves.

A target is a measured curve F(delta) (FP64, newtons, 256 samples) of a cell whose geometry and
protocol are known (mesh, cell radius, tip radius, speed, depth).  The surrogate predicts
F / (E R^2), so a candidate (E, nu, eta) predicts F = E R^2 f(features).

Surrogate-assisted fit (``fit_surrogate``), for all targets at once:
1. ``rounds`` rounds of 64 candidates per target, each round one batched inference launch over
   targets x 64 rows: round 0 samples the prior box (log10 E in [2, 4], nu in [0.3, 0.49],
   log10 eta in [0, 3]); later rounds sample a Gaussian fitted to the 8 best candidates so far
   (cross-entropy method), clipped to the box;
2. the 4 best distinct candidates per target are simulated with the batched simulator (the
   target's own mesh and protocol) and the best simulated curve is the fit -- the plain simulation
   judges it.
Fallback / baseline (``fit_nelder_mead``): Nelder-Mead on (log10 E, nu, log10 eta) with every
evaluation a simulation (all targets advance in lock-step, one batched simulation per evaluation).
"""

from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from nofluidx3d_amd.batch import CellSpec

LO = np.array([2.0, 0.3, 0.0])
HI = np.array([4.0, 0.49, 3.0])


def theta_to_spec(theta, base):
    logE, nu, logeta = theta
    return dataclasses.replace(base, E_si=10 ** float(logE), nu=float(np.clip(nu, 0.3, 0.49)),
                               viscosity_si=10 ** float(logeta))


def features(theta, bases):
    """surrogate inputs for parameter rows theta (n, 3) of cells ``bases`` (n)"""
    return np.stack([theta[:, 0], theta[:, 1], theta[:, 2], [b.tip_radius_um for b in bases],
                     [math.log10(b.speed_um_s) for b in bases], [b.radius_um for b in bases]], 1)


def curve_force(curve_norm, spec):
    """normalised curve -> newtons"""
    return curve_norm * spec.E_si * (spec.radius_um * 1e-6) ** 2


def simulate(bs, specs):
    """Curves (n, ncurve) in newtons from the batched simulator ``bs`` (same meshes as ``specs``)."""
    bs.reset(specs)
    bs.advance(bs.steps)
    cn = bs.curves_normalized()
    return np.stack([curve_force(cn[i], s) for i, s in enumerate(specs)])


def fit_surrogate(net, norm, targets, bases, bs, rounds=8, ncand=64, ntop=4, seed=0, log=None):
    """targets (n, 256) newtons; bases (n) CellSpec with the known geometry/protocol; ``bs`` a
    BatchSimulation over the targets' meshes.  Returns dict with the fitted parameters, the
    simulated curve of the fit and timings."""
    n = len(targets)
    rng = np.random.Generator(np.random.PCG64(seed))
    scale = np.array([curve_force(1.0, b) / b.E_si for b in bases])  # R^2 per target
    t0 = time.perf_counter()
    all_theta = np.zeros((n, 0, 3))
    all_err = np.zeros((n, 0))
    infer_s = 0.0
    for r in range(rounds):
        if r == 0:
            th = LO + (HI - LO) * rng.random((n, ncand, 3))
        else:
            order = np.argsort(all_err, axis=1)[:, :8]
            elite = np.take_along_axis(all_theta, order[:, :, None], 1)
            mu, sd = elite.mean(1), elite.std(1) + 0.02 * (HI - LO)
            th = np.clip(mu[:, None, :] + sd[:, None, :] * rng.normal(size=(n, ncand, 3)), LO, HI)
        rows = th.reshape(-1, 3)
        feats = features(rows, [b for b in bases for _ in range(ncand)])
        ti = time.perf_counter()
        pred = net.predict(norm.x(feats)) * norm.y_scale  # one batched inference launch chain
        infer_s += time.perf_counter() - ti
        F = pred.reshape(n, ncand, -1) * (10 ** th[:, :, :1]) * scale[:, None, None]
        err = np.linalg.norm(F - targets[:, None, :], axis=2) / np.linalg.norm(targets, axis=1)[:, None]
        all_theta = np.concatenate([all_theta, th], 1)
        all_err = np.concatenate([all_err, err], 1)
    t_sur = time.perf_counter() - t0
    # 4 best distinct candidates per target -> plain batched simulation
    order = np.argsort(all_err, axis=1)
    top = np.zeros((n, ntop, 3))
    for i in range(n):
        picked = []
        for j in order[i]:
            cand = all_theta[i, j]
            if all(np.max(np.abs(cand - p) / (HI - LO)) > 1e-3 for p in picked):
                picked.append(cand)
            if len(picked) == ntop:
                break
        top[i] = np.array(picked)
    sims = np.zeros((ntop, n, targets.shape[1]))
    t1 = time.perf_counter()
    for k in range(ntop):
        sims[k] = simulate(bs, [theta_to_spec(top[i, k], bases[i]) for i in range(n)])
        if log:
            log(f"refine {k + 1}/{ntop}: {time.perf_counter() - t1:.1f}s")
    t_sim = time.perf_counter() - t1
    serr = np.linalg.norm(sims - targets[None], axis=2) / np.linalg.norm(targets, axis=1)[None]
    best = np.argmin(serr, axis=0)
    theta = top[np.arange(n), best]
    return {"theta": theta, "curve": sims[best, np.arange(n)], "sim_err": serr[best, np.arange(n)],
            "surrogate_s": t_sur, "inference_s": infer_s, "simulation_s": t_sim,
            "total_s": time.perf_counter() - t0, "candidates_per_target": rounds * ncand,
            "simulations_per_target": ntop}


def fit_nelder_mead(targets, bases, bs, max_evals=60, log=None):
    """Simulation-only Nelder-Mead (standard coefficients 1, 2, 0.5, 0.5) on (log10 E, nu, log10 eta)
    in box-normalised coordinates, all targets in lock-step: every evaluation is one batched
    simulation of all targets."""
    n = len(targets)
    tn = np.linalg.norm(targets, axis=1)

    def evaluate(u):  # u (n, 3) in [0, 1]^3
        th = LO + (HI - LO) * np.clip(u, 0, 1)
        F = simulate(bs, [theta_to_spec(th[i], bases[i]) for i in range(n)])
        return np.linalg.norm(F - targets, axis=1) / tn

    t0 = time.perf_counter()
    simplex = np.array([[0.5, 0.5, 0.5], [0.8, 0.5, 0.5], [0.5, 0.8, 0.5], [0.5, 0.5, 0.8]])
    S = np.repeat(simplex[None], n, 0)  # (n, 4, 3)
    f = np.stack([evaluate(S[:, j]) for j in range(4)], 1)
    evals = 4
    while evals + 2 <= max_evals:
        o = np.argsort(f, 1)
        S = np.take_along_axis(S, o[:, :, None], 1)
        f = np.take_along_axis(f, o, 1)
        c = S[:, :3].mean(1)
        xr = c + (c - S[:, 3])
        fr = evaluate(xr)
        expand = fr < f[:, 0]
        # one more evaluation per target: the expansion point where it may pay, else the contraction
        x2 = np.where(expand[:, None], c + 2 * (c - S[:, 3]), c + 0.5 * (S[:, 3] - c))
        f2 = evaluate(x2)
        evals += 2
        shrink = np.zeros(n, bool)
        for i in range(n):
            if expand[i]:
                S[i, 3], f[i, 3] = (x2[i], f2[i]) if f2[i] < fr[i] else (xr[i], fr[i])
            elif fr[i] < f[i, 2]:
                S[i, 3], f[i, 3] = xr[i], fr[i]
            elif f2[i] < f[i, 3]:
                S[i, 3], f[i, 3] = x2[i], f2[i]
            else:
                shrink[i] = True
        if shrink.any() and evals + 3 <= max_evals:  # shrink towards the best vertex, re-evaluated
            S[shrink, 1:] = S[shrink, :1] + 0.5 * (S[shrink, 1:] - S[shrink, :1])
            for j in range(1, 4):
                fj = evaluate(S[:, j])
                f[shrink, j] = fj[shrink]
            evals += 3
        if log:
            log(f"nelder-mead evals {evals}: best median err {np.median(f.min(1)):.4f}")
    j = np.argmin(f, 1)
    theta = LO + (HI - LO) * np.clip(S[np.arange(n), j], 0, 1)
    return {"theta": theta, "sim_err": f[np.arange(n), j], "total_s": time.perf_counter() - t0, "evaluations": evals}
