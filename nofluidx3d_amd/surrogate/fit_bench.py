"""Fitting benchmark: surrogate-assisted fits against simulation-only Nelder-Mead fits.

Targets: ``n`` new cells from the batched workload's generator (seeds ``seed0 + i``: radius,
lattice spacing, mesh seed, tip radius, speed and the hidden E, nu, viscosity), simulated for 10k
steps, plus Gaussian noise of standard deviation 1 % of each curve's peak force (seed 5).  The
fitters see only the noisy curve and the cell's geometry and protocol.

Reported (JSON on stdout):
- ``noise_floor_rel_l2``: distance of the noisy target from its noise-free curve (no fit can do
  better against the noisy target);
- ``surrogate``: all targets fitted together by ``fit.fit_surrogate`` (``ntop`` refining
  simulations and ``corrections`` corrected rounds per target): fits/s, the batch wall time
  (the latency of every fit in the batch), the relative L2 error of the fit's simulated curve
  against the noisy target and against the noise-free curve (median, p90, p99, fraction <= 2 %),
  the same by terciles of viscosity / E, the recovered-parameter errors, and a confirmation: the
  plain simulator re-run at the fitted parameters of the first ``n_nm`` targets;
- ``nelder_mead``: the simulation-only baseline on the first ``n_nm`` targets with ``nm_evals``
  simulations per target: fits/s, errors, the median error after k evaluations and the
  evaluations it needs to reach the surrogate path's median error;
- ``variants``: other fitter settings on the same targets (errors and fits/s);
- ``single``: p50/p99 latency of one-target fits (a batch of one cell) for both methods;
- ``speedup_time_per_fit`` (both at their run lengths) and ``speedup_at_equal_median_error`` (a
  lower bound, flagged, when Nelder-Mead does not reach the surrogate path's median error).

``python -m nofluidx3d_amd.surrogate.fit_bench --weights DIR/surrogate.npz`` (on a GPU).
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hidden_blind(cells):
    """The fitters' view of the cells: geometry and protocol, placeholder material."""
    return [dataclasses.replace(c, E_si=1000.0, nu=0.4, viscosity_si=10.0) for c in cells]


def make_targets(n, seed0=7000, steps=10000, noise=0.01, noise_seed=5, tile_k=256):
    from nofluidx3d_amd.autotune import group_tetra
    from nofluidx3d_amd.batch import CellGroups, make_meshes, sample_batch
    from nofluidx3d_amd.surrogate import fit

    t0 = time.time()
    cells, _ = sample_batch(n, seed0=seed0, steps=steps)
    meshes = make_meshes(cells)
    log(f"targets: {n} meshes in {time.time() - t0:.1f}s")
    # cache-sized cell groups, each running its steps before the next (faster than one batch)
    bs = CellGroups(hidden_blind(cells), meshes, max_tetra=group_tetra(), steps=steps, ncurve=256, tile_k=tile_k)
    log(f"targets: batch built ({bs.T_real} tetra) at {time.time() - t0:.1f}s")
    clean = fit.simulate(bs, cells)
    rng = np.random.Generator(np.random.PCG64(noise_seed))
    noisy = clean + noise * np.abs(clean).max(axis=1, keepdims=True) * rng.normal(size=clean.shape)
    log(f"targets: simulated at {time.time() - t0:.1f}s")
    return cells, meshes, bs, clean, noisy


def theta_true(cells):
    return np.array([[np.log10(c.E_si), c.nu, np.log10(c.viscosity_si)] for c in cells])


def param_errors(theta, truth):
    d = np.abs(theta - truth)
    return {"log10_E_median": float(np.median(d[:, 0])), "nu_median": float(np.median(d[:, 1])),
            "log10_viscosity_median": float(np.median(d[:, 2])),
            "log10_viscosity_over_E_median": float(np.median(np.abs((theta[:, 2] - theta[:, 0]) -
                                                                     (truth[:, 2] - truth[:, 0]))))}


def err_stats(e):
    return {"median": float(np.median(e)), "p90": float(np.percentile(e, 90)), "p99": float(np.percentile(e, 99)),
            "frac_le_2pct": float(np.mean(e <= 0.02))}


def run(net, norm, args, log=log):
    """The benchmark with the settings of ``args`` (the fields of ``main``'s options); returns the
    report as a dict."""
    from nofluidx3d_amd.autotune import group_tetra
    from nofluidx3d_amd.batch import BatchSimulation, CellGroups
    from nofluidx3d_amd.surrogate import fit

    t_start = time.time()
    cells, meshes, bs, clean, noisy = make_targets(args.n, steps=args.steps)
    blind = hidden_blind(cells)
    truth = theta_true(cells)
    floor = fit.rel_err(clean, noisy)
    out = {"targets": args.n, "noise": "Gaussian, sd 1% of each curve's peak force", "steps": args.steps,
           "tetra": int(bs.T_real), "noise_floor_rel_l2": err_stats(floor)}

    r = fit.fit_surrogate(net, norm, noisy, blind, bs, ntop=args.ntop, corrections=args.corrections, log=log,
                          confirm=False)
    ratio = np.log10(np.array([c.viscosity_si / c.E_si for c in cells]))
    terc = np.quantile(ratio, [1 / 3, 2 / 3])
    by = {name: err_stats(fit.rel_err(r["curve"][m], clean[m])) for name, m in
          (("low", ratio < terc[0]), ("mid", (ratio >= terc[0]) & (ratio < terc[1])), ("high", ratio >= terc[1]))}
    out["surrogate"] = {
        "fits_per_s": args.n / r["total_s"], "batch_wall_s": r["total_s"], "surrogate_s": r["surrogate_s"],
        "inference_s": r["inference_s"], "refine_s": r["refine_s"],
        "ntop": args.ntop, "corrections": args.corrections,
        "simulations_per_target": r["simulations_per_target"], "candidates_per_target": r["candidates_per_target"],
        "rel_l2_vs_noisy_target": err_stats(r["sim_err"]), "rel_l2_vs_noise_free_curve": err_stats(
            fit.rel_err(r["curve"], clean)), "rel_l2_vs_noise_free_by_log10_viscosity_over_E_tercile": by,
        "param_errors": param_errors(r["theta"], truth)}
    log("surrogate:", json.dumps(out["surrogate"]))
    out["variants"] = []
    for v in filter(None, args.variants.split(",")):
        nt, nc = (int(q) for q in v.split(":"))
        rv = fit.fit_surrogate(net, norm, noisy, blind, bs, ntop=nt, corrections=nc, log=log, confirm=False)
        out["variants"].append({"ntop": nt, "corrections": nc, "simulations_per_target": rv["simulations_per_target"],
                                "fits_per_s": args.n / rv["total_s"],
                                "rel_l2_vs_noisy_target": err_stats(rv["sim_err"]),
                                "rel_l2_vs_noise_free_curve": err_stats(fit.rel_err(rv["curve"], clean)),
                                "param_errors": param_errors(rv["theta"], truth)})
        log("variant:", json.dumps(out["variants"][-1]))
    # confirmation: the plain simulator at the fitted parameters reproduces the fit's curve
    m = min(args.n_nm, args.n)
    bsc = CellGroups(blind[:m], meshes[:m], max_tetra=group_tetra(), steps=args.steps, ncurve=256, tile_k=256)
    conf = fit.simulate(bsc, [fit.theta_to_spec(r["theta"][i], blind[i]) for i in range(m)])
    out["surrogate"]["confirmation"] = {
        "targets": m, "max_rel_diff_confirmed_vs_fit_curve": float(np.max(fit.rel_err(conf, r["curve"][:m]))),
        "rel_l2_vs_noisy_target": err_stats(fit.rel_err(conf, noisy[:m]))}
    del bs

    nm = fit.fit_nelder_mead(noisy[:m], blind[:m], bsc, max_evals=args.nm_evals, log=log)
    reach = np.argmax(nm["history"] <= 0.02, axis=0).astype(float) + 1
    reach[~(nm["history"] <= 0.02).any(axis=0)] = np.nan
    out["nelder_mead"] = {
        "targets": m, "fits_per_s": m / nm["total_s"], "batch_wall_s": nm["total_s"], "evaluations": nm["evaluations"],
        "rel_l2": err_stats(nm["sim_err"]), "param_errors": param_errors(nm["theta"], truth[:m]),
        "frac_reaching_2pct": float(np.mean(np.isfinite(reach))),
        "evals_to_2pct_median_of_reaching": float(np.nanmedian(reach)) if np.isfinite(reach).any() else None,
        "median_err_after_evals": {str(k): float(np.median(nm["history"][k - 1]))
                                   for k in (4, 10, 20, 40, 60, 80, 100, 150, 200, 250, 300, 400)
                                   if k <= len(nm["history"])}}
    # evaluations Nelder-Mead needs for the surrogate path's median error (vs the noisy target)
    goal = float(np.median(r["sim_err"][:m]))
    med = np.median(nm["history"], axis=1)
    hit = np.nonzero(med <= goal)[0]
    out["nelder_mead"]["evals_to_surrogate_median_error"] = int(hit[0] + 1) if len(hit) else None
    out["nelder_mead"]["surrogate_median_error_same_targets"] = goal
    log("nelder-mead:", json.dumps(out["nelder_mead"]))
    del bsc
    out["speedup_time_per_fit"] = (nm["total_s"] / m) / (r["total_s"] / args.n)
    ev = out["nelder_mead"]["evals_to_surrogate_median_error"]
    # when Nelder-Mead does not reach the surrogate path's median error within its evaluations,
    # the ratio at its last evaluation is a lower bound
    out["speedup_at_equal_median_error"] = (nm["total_s"] / m) * ((ev or nm["evaluations"]) / nm["evaluations"]) / (
        r["total_s"] / args.n)
    out["speedup_at_equal_median_error_is_lower_bound"] = ev is None
    # the same comparison in simulations per target (independent of how well each method's batch
    # fills the GPU: Nelder-Mead runs on n_nm targets, the surrogate path on all n)
    out["simulations_ratio_at_equal_median_error"] = (ev or nm["evaluations"]) / r["simulations_per_target"]

    lat_s, lat_n = [], []
    for i in range(args.single):
        b1 = BatchSimulation(blind[i:i + 1], meshes[i:i + 1], steps=args.steps, ncurve=256, tile_k=256)
        t0 = time.perf_counter()
        fit.fit_surrogate(net, norm, noisy[i:i + 1], blind[i:i + 1], b1, ntop=args.ntop,
                          corrections=args.corrections, confirm=False)
        lat_s.append(time.perf_counter() - t0)
        if i < args.single_nm:
            t0 = time.perf_counter()
            fit.fit_nelder_mead(noisy[i:i + 1], blind[i:i + 1], b1, max_evals=args.nm_evals)
            lat_n.append(time.perf_counter() - t0)
        del b1
    if lat_s:
        out["single"] = {"surrogate_p50_s": float(np.percentile(lat_s, 50)),
                         "surrogate_p99_s": float(np.percentile(lat_s, 99)), "samples": [len(lat_s), len(lat_n)]}
        if lat_n:
            out["single"].update({"nelder_mead_p50_s": float(np.percentile(lat_n, 50)),
                                  "nelder_mead_p99_s": float(np.percentile(lat_n, 99))})
    out["wall_s"] = time.time() - t_start
    return out


def options(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", required=True)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--n-nm", type=int, default=512)
    ap.add_argument("--nm-evals", type=int, default=60)
    ap.add_argument("--ntop", type=int, default=1)
    ap.add_argument("--corrections", type=int, default=1)
    ap.add_argument("--variants", default="",
                    help="more fitter settings on the same targets, 'ntop:corrections,...' (reported, not timed "
                         "against Nelder-Mead)")
    ap.add_argument("--single", type=int, default=16, help="one-target surrogate fits timed for p50/p99")
    ap.add_argument("--single-nm", type=int, default=2, help="one-target Nelder-Mead fits timed")
    ap.add_argument("--steps", type=int, default=10000)
    return ap.parse_args(argv)


def main(argv=None):
    from nofluidx3d_amd.surrogate import train

    args = options(argv)
    net, norm = train.load(args.weights)
    print(json.dumps(run(net, norm, args)), flush=True)


if __name__ == "__main__":
    main()
