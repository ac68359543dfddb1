"""Training data from the batched simulator and training of the curve surrogate.

Data: a pool of cell meshes (the batched workload's cells, seeds 1000+i) is simulated chunk after
chunk with new physical parameters drawn from the batched workload's distribution (E log-uniform
0.1-10 kPa, viscosity log-uniform 1-1000 Pa s, nu U(0.3, 0.49), tip radius U(1, 5) um, speed
log-uniform 0.5-50 um/s; per-curve seeds ``seed0 + i``), 10k steps each, one 256-point curve per
cell, normalised by E R^2.  Held-out curves use disjoint seeds.

Model: MLP 6 -> 512 x 4 (GELU) -> 256, inputs standardised (log10 E, nu, log10 eta, tip radius,
log10 speed, cell radius), targets scaled by one global constant; Adam (lr 1e-3, batch 4096).
Error metric: relative L2 error of each held-out curve, ||pred - sim|| / ||sim||.

``python -m nofluidx3d_amd.surrogate.train --out DIR`` (on a GPU) writes ``DIR/surrogate.npz``
(weights + normalisation), ``DIR/heldout.npz`` (held-out inputs and curves) and ``DIR/report.json``.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from nofluidx3d_amd.batch import CellSpec

SIZES = (6, 512, 512, 512, 512, 256)


def sample_physics(pool, n, seed0, steps, max_relax=0.1):
    """Cell i keeps the mesh of pool[i % len(pool)] and draws its physics from seed seed0 + i."""
    out = []
    for i in range(n):
        base = pool[i % len(pool)]
        rng = np.random.Generator(np.random.PCG64(seed0 + i))
        while True:
            E = 10 ** rng.uniform(2.0, 4.0)
            eta = 10 ** rng.uniform(0.0, 3.0)
            nu = rng.uniform(0.3, 0.49)
            v = 10 ** rng.uniform(math.log10(0.5), math.log10(50.0))
            tip = rng.uniform(1.0, 5.0)
            s = CellSpec(base.radius_um, base.spacing_um, base.seed, E, nu, eta, tip, v)
            if s.relaxation_per_step(steps) <= max_relax:
                break
        out.append(s)
    return out


class Normaliser:
    def __init__(self, x_mean, x_std, y_scale):
        self.x_mean, self.x_std, self.y_scale = np.asarray(x_mean), np.asarray(x_std), float(y_scale)

    def x(self, X):
        return (np.asarray(X) - self.x_mean) / self.x_std


def generate(bs, pool, n_chunks, seed0, log=print):
    """Run ``n_chunks`` chunks of len(pool) cells; returns (features, normalised curves)."""
    X, Y = [], []
    for c in range(n_chunks):
        t0 = time.time()
        specs = sample_physics(pool, len(pool), seed0 + c * len(pool), bs.steps)
        bs.reset(specs)
        bs.advance(bs.steps)
        Y.append(bs.curves_normalized().astype(np.float32))
        X.append(bs.features().astype(np.float32))
        log(f"chunk {c + 1}/{n_chunks}: {len(pool)} curves in {time.time() - t0:.1f}s", flush=True)
    X, Y = np.concatenate(X), np.concatenate(Y)
    ok = np.isfinite(Y).all(1)
    return X[ok], Y[ok]


def train(X, Y, iters=4000, batch=4096, seed=0, log=print, Xv=None, Yv=None, every=500):
    import torch

    from nofluidx3d_amd.surrogate.mlp import MLP

    x_mean, x_std = X.mean(0), X.std(0) + 1e-12
    y_scale = float(Y.std())
    norm = Normaliser(x_mean, x_std, y_scale)
    net = MLP(SIZES, seed=seed)
    Xn = norm.x(X).astype(np.float32)
    Yn = (Y / y_scale).astype(np.float32)
    Yd = torch.from_numpy(Yn).to(net.dev)
    rng = np.random.Generator(np.random.PCG64(seed))
    B = min(batch, -(-len(X) // 128) * 128)
    hist = []
    t0 = time.time()
    for it in range(iters):
        idx = rng.integers(0, len(X), B)
        net.set_input(Xn[idx], train=True)
        tgt = Yd[torch.from_numpy(idx).to(net.dev)]
        net.train_step(B, tgt)
        if (it + 1) % every == 0 or it == iters - 1:
            net.sync()
            rec = {"iter": it + 1, "train_mse": net.loss_dev.item(), "seconds": time.time() - t0}
            if Xv is not None:
                rec["heldout_rel_l2_median"] = float(np.median(curve_errors(net, norm, Xv, Yv)))
            hist.append(rec)
            log(rec, flush=True)
    return net, norm, hist


def curve_errors(net, norm, X, Y):
    pred = net.predict(norm.x(X)) * norm.y_scale
    return np.linalg.norm(pred - Y, axis=1) / np.linalg.norm(Y, axis=1)


def save(path, net, norm):
    st = net.state()
    np.savez(path, x_mean=norm.x_mean, x_std=norm.x_std, y_scale=norm.y_scale, sizes=np.array(net.sizes), **st)


def load(path):
    from nofluidx3d_amd.surrogate.mlp import MLP

    d = np.load(path)
    net = MLP(tuple(int(v) for v in d["sizes"]), seed=0)
    net.load_state({k: d[k] for k in d.files if k[0] in "Wb" and k[1:].isdigit()})
    return net, Normaliser(d["x_mean"], d["x_std"], float(d["y_scale"]))


def main(argv=None):
    from nofluidx3d_amd.batch import BatchSimulation, make_meshes, sample_batch

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--chunks", type=int, default=40)
    ap.add_argument("--heldout-chunks", type=int, default=2)
    ap.add_argument("--pool", type=int, default=512)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--tile-k", type=int, default=256)
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    log = lambda *a, **k: print(*a, **k, file=sys.stderr)
    t0 = time.time()
    pool, _ = sample_batch(args.pool, seed0=1000, steps=args.steps)
    meshes = make_meshes(pool)
    bs = BatchSimulation(pool, meshes, steps=args.steps, ncurve=256, tile_k=args.tile_k)
    Xt, Yt = generate(bs, pool, args.chunks, seed0=100000, log=log)
    Xv, Yv = generate(bs, pool, args.heldout_chunks, seed0=900000, log=log)
    tgen = time.time() - t0
    np.savez(os.path.join(args.out, "heldout.npz"), X=Xv, Y=Yv)
    net, norm, hist = train(Xt, Yt, iters=args.iters, log=log, Xv=Xv, Yv=Yv)
    err = curve_errors(net, norm, Xv, Yv)
    save(os.path.join(args.out, "surrogate.npz"), net, norm)
    rep = {
        "train_curves": int(len(Xt)),
        "heldout_curves": int(len(Xv)),
        "data_generation_s": tgen,
        "tetra_per_chunk": int(bs.T_real),
        "iters": args.iters,
        "history": hist,
        "heldout_rel_l2": {"mean": float(err.mean()), "median": float(np.median(err)),
                           "p95": float(np.percentile(err, 95)), "max": float(err.max()),
                           "frac_le_2pct": float((err <= 0.02).mean())},
    }
    json.dump(rep, open(os.path.join(args.out, "report.json"), "w"), indent=1)
    print(json.dumps(rep["heldout_rel_l2"]))


if __name__ == "__main__":
    main()
