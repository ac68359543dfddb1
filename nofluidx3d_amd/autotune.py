"""Launch-shape autotuner for the tiled force pass (single cells and batches).

Candidates are timed on the device with HIP events over a few steps (``nfx_time_passes`` /
``nfx_batch_time_passes``); the fastest wins and is cached per problem signature in a JSON file
(``$NFX_TUNE_CACHE`` or ``~/.cache/nofluidx3d_amd/tune.json``).  Every candidate computes the same
physics; tile size only changes the floating-point summation order, by rounding.

Tunables:
- ``tile_k``: tetra per tile = threads per workgroup of the force pass (256 or 512);
- ``graph_steps``: steps per captured hipGraph replay block (the library captures at most 100).
"""

from __future__ import annotations

import json
import os
import time

TILE_CANDIDATES = (256, 512)


def _cache_path():
    return os.environ.get("NFX_TUNE_CACHE", os.path.expanduser("~/.cache/nofluidx3d_amd/tune.json"))


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    d = _load()
    d[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as f:
            json.dump(d, f, indent=1)
        os.replace(path + ".tmp", path)
    except OSError:
        pass


def tune_batch(specs, meshes, steps, ncurve=256, nrep=10, use_cache=True, log=None):
    """Pick ``tile_k`` for a batch by timing one force+node+curve pass per candidate.
    Returns (best_tile_k, {tile_k: ms_per_step})."""
    from nofluidx3d_amd.batch import BatchSimulation

    tetra = sum(len(m[1]) for m in meshes)
    key = f"batch:{len(specs)}:{tetra}:{specs[0].viscosity_si is not None}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            return hit["tile_k"], {int(k): v for k, v in hit["ms"].items()}
    times = {}
    for K in TILE_CANDIDATES:
        try:
            b = BatchSimulation(specs, meshes, steps=steps, ncurve=ncurve, tile_k=K)
        except Exception as e:  # a tile with more nodes than threads: candidate not applicable
            if log:
                log(f"tile_k {K}: {e}")
            continue
        b.time_passes(2)
        f, n, c = b.time_passes(nrep)
        times[K] = f + n + c
        if log:
            log(f"tile_k {K}: force {f:.4f} ms, node {n:.4f} ms, curve {c:.4f} ms")
        del b
    best = min(times, key=times.get)
    _store(key, {"tile_k": best, "ms": times, "when": time.time()})
    return best, times


def tune_simulation(sim, nrep=20, log=None):
    """Pick ``tile_k`` for a single-cell Simulation with the tiled scatter (mutates ``sim.tile_k``)."""
    import ctypes

    import torch

    from nofluidx3d_amd import _lib, openCL

    if sim.scatter != "tiled":
        return sim.tile_k, {}
    saved_time = sim.time
    sim.sync_forms()
    openCL.getCommandQueue().finish()
    snapshot = [t.clone() for t in sim.state_tensors()]  # timing runs real steps: restore afterwards
    times = {}
    for K in TILE_CANDIDATES:
        sim.tile_k = K
        sim._tiles = None
        sim._plan = None
        p = sim.plan()
        if p is None or sim.scatter != "tiled":
            sim.scatter = "tiled"
            continue
        mf, mn = ctypes.c_double(), ctypes.c_double()
        _lib.call("nfx_time_passes", openCL.handle(), p, sim.time, nrep, ctypes.byref(mf), ctypes.byref(mn))
        sim.time += nrep
        times[K] = mf.value + mn.value
        if log:
            log(f"tile_k {K}: force {mf.value:.4f} ms, node {mn.value:.4f} ms")
    openCL.getCommandQueue().finish()
    for t, s0 in zip(sim.state_tensors(), snapshot):
        t.copy_(s0)
    torch.cuda.synchronize()
    sim.forms_restored()
    sim.time = saved_time
    best = min(times, key=times.get)
    sim.tile_k = best
    sim._tiles = None
    sim._plan = None
    return best, times
