"""GPU tests of the surrogate pipeline end to end at small size: curves from the batched simulator,
training with the MFMA kernels, held-out error, the exact E-viscosity symmetry the fitter and the
training augmentation rely on, and the surrogate-assisted fit against the simulation-only
Nelder-Mead at a fixed simulation budget."""

import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pipeline():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd.batch import BatchSimulation, CellSpec, make_meshes
    from nofluidx3d_amd.surrogate import train

    pool = [CellSpec(3.0 + 0.2 * i, 0.75, 4000 + i, 1000.0, 0.4, 100.0, 1.5, 5.0) for i in range(16)]
    meshes = make_meshes(pool, workers=1, cache=None)
    bs = BatchSimulation(pool, meshes, steps=600, ncurve=256, tile_k=256)
    Xt, Yt = train.generate(bs, pool, 12, seed0=100, log=lambda *a, **k: None)
    Xv, Yv = train.generate(bs, pool, 1, seed0=9000, log=lambda *a, **k: None)
    net, norm, hist = train.train(Xt, Yt, iters=600, batch=256, log=lambda *a, **k: None, Xv=Xv, Yv=Yv,
                                  every=100)
    return pool, bs, net, norm, hist, (Xv, Yv)


def test_training_converges_and_generalises(pipeline):
    pool, bs, net, norm, hist, (Xv, Yv) = pipeline
    from nofluidx3d_amd.surrogate.train import curve_errors

    losses = [h["train_loss"] for h in hist if "train_loss" in h]
    assert losses[-1] < 0.5 * losses[0]
    err = curve_errors(net, norm, Xv, Yv)
    assert np.median(err) < 0.15, np.median(err)


def test_e_viscosity_scaling_is_exact(pipeline):
    """(E, viscosity) -> (a E, a viscosity): identical run in simulation units, curve a times."""
    pool, bs, *_ = pipeline
    from nofluidx3d_amd.surrogate import fit

    a = 3.7
    specs = [dataclasses.replace(s, E_si=500.0 + 40 * i, viscosity_si=20.0 + i) for i, s in enumerate(pool)]
    scaled = [dataclasses.replace(s, E_si=a * s.E_si, viscosity_si=a * s.viscosity_si) for s in specs]
    F1 = fit.simulate(bs, specs)
    n1 = bs.curves_normalized()
    F2 = fit.simulate(bs, scaled)
    n2 = bs.curves_normalized()
    assert np.max(np.abs(n1)) > 0
    assert np.max(np.abs(n2 - n1)) <= 1e-6 * np.max(np.abs(n1))
    assert np.max(np.abs(F2 - a * F1)) <= 1e-6 * np.max(np.abs(a * F1))


def test_surrogate_fit_beats_simulation_only_at_equal_budget(pipeline):
    """Surrogate screen + 1 refining simulation + 1 corrected round + 1 confirming simulation per
    target against Nelder-Mead with 6 simulations per target (the initial simplex and one step)."""
    pool, bs, net, norm, hist, _ = pipeline
    from nofluidx3d_amd.surrogate import fit, train

    truth = train.sample_physics(pool, len(pool), 777, bs.steps)
    targets = fit.simulate(bs, truth)
    blind = [dataclasses.replace(c, E_si=1000.0, nu=0.4, viscosity_si=10.0) for c in truth]
    sur = fit.fit_surrogate(net, norm, targets, blind, bs, rounds=6, ntop=1, corrections=1)
    nm = fit.fit_nelder_mead(targets, blind, bs, max_evals=6)
    assert sur["simulations_per_target"] == 3
    assert np.median(sur["sim_err"]) < np.median(nm["sim_err"])
    # the confirming simulation reproduces the refined (amplitude-corrected) curve wherever the
    # corrected parameters stayed inside the box
    same = np.isclose(sur["sim_err"], sur["predicted_refined_err"], rtol=1e-4, atol=1e-6)
    assert np.mean(same) >= 0.75, (sur["sim_err"], sur["predicted_refined_err"])
