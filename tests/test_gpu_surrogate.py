"""GPU test of the surrogate pipeline end to end at small size: curves from the batched simulator,
training with the MFMA kernels, held-out error, and the surrogate-assisted fit against the
simulation-only Nelder-Mead at a fixed simulation budget."""

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
                                  every=200)
    return pool, bs, net, norm, hist, (Xv, Yv)


def test_training_converges_and_generalises(pipeline):
    pool, bs, net, norm, hist, (Xv, Yv) = pipeline
    from nofluidx3d_amd.surrogate.train import curve_errors

    assert hist[-1]["train_mse"] < 0.05 * 1.0  # targets are scaled to unit standard deviation
    err = curve_errors(net, norm, Xv, Yv)
    assert np.median(err) < 0.15, np.median(err)


def test_surrogate_fit_beats_simulation_only_at_equal_budget(pipeline):
    """4 simulations per target after the surrogate screen vs Nelder-Mead with the same number of
    simulations (6: the initial simplex and one step)."""
    pool, bs, net, norm, hist, _ = pipeline
    from nofluidx3d_amd.surrogate import fit, train

    truth = train.sample_physics(pool, len(pool), 777, bs.steps)
    targets = fit.simulate(bs, truth)
    sur = fit.fit_surrogate(net, norm, targets, truth, bs, rounds=6)
    nm = fit.fit_nelder_mead(targets, truth, bs, max_evals=6)
    assert np.median(sur["sim_err"]) < np.median(nm["sim_err"])
    assert sur["simulations_per_target"] == 4
