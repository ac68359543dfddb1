"""Training data from the batched simulator and training of the curve surrogate.

Data.  A pool of cell meshes (the batched workload's cells, seeds 1000+i) is simulated chunk after
chunk with new physical parameters drawn from the batched workload's distribution (E log-uniform
0.1-10 kPa, viscosity log-uniform 1-1000 Pa s, nu U(0.3, 0.49), tip radius U(1, 5) um, speed
log-uniform 0.5-50 um/s; per-curve seeds ``seed0 + i``), 10k steps each, one 256-point curve per
cell, normalised by E R^2.  Two held-out sets: new physics on the pool's meshes (disjoint seeds),
and new cells (new meshes, seeds 5000+i) with new physics.

Model.  MLP 6 -> 512 x 4 (GELU) -> 256 on ``batch.surrogate_inputs`` (standardised: log10
viscosity / E, nu, log10 tip radius, log10 speed, log10 cell radius, log10 lattice spacing),
targets scaled by one global constant; Adam (lr 1e-3, batch 4096).
- Loss: the squared relative L2 error of each curve, i.e. the MSE with row weights
  1 / mean(target_row^2) (floored at 1e-3 of the median), so that small-force curves count as much
  as large ones.
- An exact symmetry of the simulator is built into the inputs: scaling E and the viscosity by one
  factor leaves every simulation-unit quantity of a cell unchanged (p0 scales with E, the flow
  parameter 3 eta / p0 and the time scale do not change), so the normalised curve depends on their
  ratio only, which is one input.
- Early stopping: 5 % of the training curves are a validation split; the weights with the lowest
  median validation error are kept.
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
# the network's inputs, in the order of batch.surrogate_inputs (stored with the weights)
INPUTS = ("log10_viscosity_over_E", "nu", "log10_tip_radius_um", "log10_speed_um_s", "log10_radius_um",
          "log10_spacing_um")
LOG_E = (2.0, 4.0)  # sampled range of log10 E [Pa]
LOG_ETA = (0.0, 3.0)  # sampled range of log10 viscosity [Pa s]


def sample_physics(pool, n, seed0, steps, max_relax=0.1):
    """Cell i keeps the mesh of pool[i % len(pool)] and draws its physics from seed seed0 + i."""
    out = []
    for i in range(n):
        base = pool[i % len(pool)]
        rng = np.random.Generator(np.random.PCG64(seed0 + i))
        while True:
            E = 10 ** rng.uniform(*LOG_E)
            eta = 10 ** rng.uniform(*LOG_ETA)
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
    """Run ``n_chunks`` chunks of len(pool) cells; returns (inputs, normalised curves)."""
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


def row_weights(Y):
    """1 / mean(y^2) per curve (floored at 1e-3 of the median), normalised to mean 1."""
    ms = np.mean(np.asarray(Y, np.float64) ** 2, axis=1)
    w = 1.0 / np.maximum(ms, 1e-3 * np.median(ms))
    return (w / w.mean()).astype(np.float32)


def train(X, Y, iters=4000, batch=4096, seed=0, log=print, Xv=None, Yv=None, every=250, relative=True,
          val_frac=0.05):
    import torch

    from nofluidx3d_amd.surrogate.mlp import MLP

    rng = np.random.Generator(np.random.PCG64(seed))
    perm = rng.permutation(len(X))
    nval = max(1, int(val_frac * len(X))) if len(X) > 20 else 0
    val, tr = perm[:nval], perm[nval:]
    Xtr, Ytr = X[tr], Y[tr]
    # FP64 statistics; an input that is constant over the data (e.g. one lattice spacing) keeps
    # scale 1, so rounding noise around its mean is not blown up
    x_mean, x_std = X.astype(np.float64).mean(0), X.astype(np.float64).std(0)
    x_std = np.where(x_std > 1e-6, x_std, 1.0)
    y_scale = float(Y.std())
    norm = Normaliser(x_mean, x_std, y_scale)
    net = MLP(SIZES, seed=seed)
    Yd = torch.from_numpy((Ytr / y_scale).astype(np.float32)).to(net.dev)
    Wd = torch.from_numpy(row_weights(Ytr)).to(net.dev) if relative else None
    B = min(batch, -(-len(Xtr) // 128) * 128)
    hist = []
    best = (np.inf, None, 0)
    t0 = time.time()
    for it in range(iters):
        idx = rng.integers(0, len(Xtr), B)
        net.set_input(norm.x(Xtr[idx]), train=True)
        di = torch.from_numpy(idx).to(net.dev)
        net.train_step(B, Yd[di], None if Wd is None else Wd[di])
        if (it + 1) % every == 0 or it == iters - 1:
            net.sync()
            rec = {"iter": it + 1, "train_loss": net.loss_dev.item(), "seconds": time.time() - t0}
            if nval:
                ev = float(np.median(curve_errors(net, norm, X[val], Y[val])))
                rec["val_rel_l2_median"] = ev
                if ev < best[0]:
                    best = (ev, net.state(), it + 1)
            if Xv is not None:
                rec["heldout_rel_l2_median"] = float(np.median(curve_errors(net, norm, Xv, Yv)))
            hist.append(rec)
            log(rec, flush=True)
    if best[1] is not None:
        net.load_state(best[1])
        hist.append({"kept_iter": best[2], "val_rel_l2_median": best[0]})
    return net, norm, hist


def curve_errors(net, norm, X, Y):
    pred = net.predict(norm.x(X)) * norm.y_scale
    return np.linalg.norm(pred - Y, axis=1) / np.linalg.norm(Y, axis=1)


def error_summary(err):
    return {"n": int(len(err)), "mean": float(err.mean()), "median": float(np.median(err)),
            "p90": float(np.percentile(err, 90)), "p95": float(np.percentile(err, 95)), "max": float(err.max()),
            "frac_le_2pct": float((err <= 0.02).mean()), "frac_le_5pct": float((err <= 0.05).mean())}


def save(path, net, norm):
    st = net.state()
    np.savez(path, x_mean=norm.x_mean, x_std=norm.x_std, y_scale=norm.y_scale, sizes=np.array(net.sizes),
             inputs=np.array(INPUTS), **st)


def load(path):
    """Network and normalisation saved by ``save``; weights trained on another input layout are
    refused (their inputs would be misread)."""
    from nofluidx3d_amd.surrogate.mlp import MLP

    d = np.load(path)
    got = tuple(str(v) for v in d["inputs"]) if "inputs" in d.files else ("(unnamed, before the ratio input)",)
    if got != INPUTS:
        raise ValueError(f"{path}: weights trained on inputs {got}, this build feeds {INPUTS}")
    net = MLP(tuple(int(v) for v in d["sizes"]), seed=0)
    net.load_state({k: d[k] for k in d.files if k[0] in "Wb" and k[1:].isdigit()})
    return net, Normaliser(d["x_mean"], d["x_std"], float(d["y_scale"]))


def _population(pool_size, mesh_seed0, steps, tile_k):
    from nofluidx3d_amd.autotune import group_tetra
    from nofluidx3d_amd.batch import CellGroups, make_meshes, sample_batch

    pool, _ = sample_batch(pool_size, seed0=mesh_seed0, steps=steps)
    return pool, CellGroups(pool, make_meshes(pool), max_tetra=group_tetra(), steps=steps, ncurve=256, tile_k=tile_k)


def write_part(prefix, X, Y, meta):
    """One slice of training data: inputs (float32), normalised curves (float16: 11 bits, far below
    the per-cent errors the network is judged at) and what produced them."""
    np.save(prefix + "_X.npy", X.astype(np.float32))
    np.save(prefix + "_Y.npy", Y.astype(np.float16))
    json.dump(meta, open(prefix + "_meta.json", "w"), indent=1)


def read_parts(prefixes):
    X = np.concatenate([np.load(p + "_X.npy") for p in prefixes])
    Y = np.concatenate([np.load(p + "_Y.npy").astype(np.float32) for p in prefixes])
    return X, Y


def read_metas(prefixes):
    out = []
    for p in prefixes:
        try:
            out.append(json.load(open(p + "_meta.json")))
        except (OSError, ValueError):
            out.append(None)
    return out


def main(argv=None):
    """``all`` (default): generate, train and evaluate in one run.  For data sets larger than one
    run can generate, ``generate`` writes slices (``--first-chunk``/``--chunks`` of a pool) and
    ``fit`` trains on slices and evaluates on a held-out slice."""
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", nargs="?", default="all", choices=["all", "generate", "fit"])
    ap.add_argument("--out", required=True, help="output directory (all, fit) or file prefix (generate)")
    ap.add_argument("--chunks", type=int, default=130)
    ap.add_argument("--first-chunk", type=int, default=0)
    ap.add_argument("--seed0", type=int, default=100000, help="physics seeds: seed0 + chunk * pool + i")
    ap.add_argument("--mesh-seed0", type=int, default=1000, help="meshes: the batched workload's cells from here")
    ap.add_argument("--heldout-chunks", type=int, default=6)
    ap.add_argument("--newmesh-chunks", type=int, default=4)
    ap.add_argument("--pool", type=int, default=512)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--iters", type=int, default=12000)
    ap.add_argument("--tile-k", type=int, default=256)
    ap.add_argument("--train", nargs="*", default=[], help="fit: training slice prefixes")
    ap.add_argument("--heldout", nargs="*", default=[], help="fit: held-out slice prefixes (new cells)")
    ap.add_argument("--heldout-physics", nargs="*", default=[],
                    help="fit: held-out slice prefixes (new physics on the training pool's meshes)")
    args = ap.parse_args(argv)
    log = lambda *a, **k: print(*a, **k, file=sys.stderr)  # noqa: E731
    t0 = time.time()
    if args.mode == "generate":
        pool, bs = _population(args.pool, args.mesh_seed0, args.steps, args.tile_k)
        X, Y = generate(bs, pool, args.chunks, seed0=args.seed0 + args.first_chunk * len(pool), log=log)
        write_part(args.out, X, Y, {"pool": args.pool, "mesh_seed0": args.mesh_seed0, "seed0": args.seed0,
                                    "first_chunk": args.first_chunk, "chunks": args.chunks, "steps": args.steps,
                                    "curves": int(len(X)), "seconds": time.time() - t0})
        print(json.dumps({"curves": int(len(X)), "seconds": time.time() - t0}))
        return
    os.makedirs(args.out, exist_ok=True)
    if args.mode == "fit":
        Xt, Yt = read_parts(args.train)
        Xn, Yn = read_parts(args.heldout)
        split = len(Xn) // 2  # the first half reports progress during training, all of it is scored
        t1 = time.time()
        net, norm, hist = train(Xt, Yt, iters=args.iters, log=log, Xv=Xn[:split], Yv=Yn[:split])
        save(os.path.join(args.out, "surrogate.npz"), net, norm)
        metas = read_metas(args.train + args.heldout + args.heldout_physics)
        rep = {"train_curves": int(len(Xt)), "train_slices": args.train, "heldout_slices": args.heldout,
               "heldout_physics_slices": args.heldout_physics,
               "data_generation_s": float(sum(m.get("seconds", 0.0) for m in metas if m)),
               "slices": metas, "training_s": time.time() - t1, "iters": args.iters, "history": hist,
               "heldout_cells_rel_l2": error_summary(curve_errors(net, norm, Xn, Yn))}
        if args.heldout_physics:
            Xp, Yp = read_parts(args.heldout_physics)
            rep["heldout_physics_rel_l2"] = error_summary(curve_errors(net, norm, Xp, Yp))
        json.dump(rep, open(os.path.join(args.out, "report.json"), "w"), indent=1)
        print(json.dumps({k: rep[k] for k in ("train_curves", "heldout_cells_rel_l2", "heldout_physics_rel_l2")
                          if k in rep}))
        return
    pool, bs = _population(args.pool, args.mesh_seed0, args.steps, args.tile_k)
    Xt, Yt = generate(bs, pool, args.chunks, seed0=args.seed0, log=log)
    Xv, Yv = generate(bs, pool, args.heldout_chunks, seed0=900000, log=log)
    del bs
    pool2, bs2 = _population(args.pool, 5000, args.steps, args.tile_k)
    Xn, Yn = generate(bs2, pool2, args.newmesh_chunks, seed0=950000, log=log)
    del bs2
    tgen = time.time() - t0
    np.savez(os.path.join(args.out, "heldout.npz"), X=Xv, Y=Yv, Xn=Xn, Yn=Yn)
    t1 = time.time()
    net, norm, hist = train(Xt, Yt, iters=args.iters, log=log, Xv=Xv, Yv=Yv)
    ttrain = time.time() - t1
    save(os.path.join(args.out, "surrogate.npz"), net, norm)
    rep = {
        "train_curves": int(len(Xt)),
        "data_generation_s": tgen,
        "training_s": ttrain,
        "iters": args.iters,
        "history": hist,
        "heldout_physics_rel_l2": error_summary(curve_errors(net, norm, Xv, Yv)),
        "heldout_cells_rel_l2": error_summary(curve_errors(net, norm, Xn, Yn)),
    }
    json.dump(rep, open(os.path.join(args.out, "report.json"), "w"), indent=1)
    print(json.dumps({k: rep[k] for k in ("heldout_physics_rel_l2", "heldout_cells_rel_l2")}))


if __name__ == "__main__":
    main()
