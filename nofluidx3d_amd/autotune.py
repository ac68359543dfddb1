"""Launch-shape autotuner for the tiled force pass (single cells and batches).

Candidates are timed over graph-replayed steps, launch costs included (a single cell: each tile
size, cache policy and step form; a batch: each tile size and cache policy, and cell groups); the
fastest wins; batch choices are cached per problem signature in a JSON file
(``$NFX_TUNE_CACHE`` or ``~/.cache/nofluidx3d_amd/tune.json``).  Every candidate computes the same
physics; tile size only changes the floating-point summation order, by rounding.

Tunables:
- ``tile_k``: tetra per tile = threads per workgroup of the force pass (128, 256 or 512; 128 gives a
  small mesh enough workgroups to fill the chip, at the cost of more tile-node halo);
- ``stream_nt``: cache policy of the force pass's per-tetra streams (state, tile tables): 1 =
  nontemporal, 0 = default.  A state that fits MI355X's 256 MB memory-side cache is served from it
  from step to step under the default policy; a larger one streams faster past it;
- the step form (single cells): node-fused (the node update of each step rides in the next force
  pass: one launch per step) or two launches per step;
- ``graph_steps``: steps per captured hipGraph replay block (the library captures at most 100);
- ``group_tetra`` (batches): run the cells as consecutive groups of at most this many tetra, each
  group all its steps before the next (``batch.CellGroups``), so that a group's working set stays
  in the memory-side cache from step to step; 0 = one batch.
"""

from __future__ import annotations

import json
import os
import time

TILE_CANDIDATES = (256, 512)
# a single cell also tries 128-tetra tiles when 512-tetra tiles would leave the chip's 256 CUs idle
SMALL_MESH_TETRA = 256 * 512
STREAM_CANDIDATES = (0, 1)
MEMORY_SIDE_CACHE_BYTES = 256 << 20
STATE_BYTES_PER_TETRA = 48  # force pass streams: 9-float state, plastic strain or volume, 4 slots
# a batch's per-step working set per tetra: symmetric state 24 B, corner records 8 B, tile-node ids,
# partial slots and positions (0.42 tile-nodes per tetra) and the nodes' state (about 0.16 nodes per
# tetra), an estimate the autotuner checks by timing
RESIDENT_BYTES_PER_TETRA = 56


def default_stream_nt(tetra):
    """the cache policy to use untuned: nontemporal once the streamed state exceeds the cache"""
    return int(STATE_BYTES_PER_TETRA * int(tetra) > MEMORY_SIDE_CACHE_BYTES)


# the node-fused step (one launch per step) wins where launches dominate a step; on a mesh of
# millions of tetra its force pass, which then also updates every node it touches (~2.9 times per
# node), is longer than the two passes apart.  Measured on MI355X (NH cells, best tile shape of each
# form, us per graph-replayed step, fused against two launches): 25k tetra 5.5 / 7.2, 106k 6.3 / 7.9,
# 211k 7.5 / 8.8, 526k 11.7 / 12.5, 878k 16.6 / 16.5; the 3.4M-tetra viscoplastic cell 58 / 50
NODE_FUSED_MAX_TETRA = 600_000


def default_node_fused(tetra):
    """the step form to use untuned: node-fused for meshes of at most NODE_FUSED_MAX_TETRA tetra"""
    return int(tetra) <= NODE_FUSED_MAX_TETRA


def group_tetra():
    """tetra per cell group whose per-step working set fits the memory-side cache"""
    return MEMORY_SIDE_CACHE_BYTES // RESIDENT_BYTES_PER_TETRA


def _label(K, nt):
    return f"{K}{'nt' if nt else ''}"


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


def _time_steps(b, nsteps):
    """ms per step of ``nsteps`` graph-replayed steps of a batch, or of cell groups summed over the
    groups: each group warms up (100 steps, capturing its graph) right before its own timed steps,
    as it runs all its steps in one call in use (the library caches one batch graph at a time)"""
    total = 0.0
    for part in getattr(b, "groups", [b]):
        part.advance(100)
        part.sync()
        t0 = time.perf_counter()
        part.advance(nsteps)
        part.sync()
        total += (time.perf_counter() - t0) / nsteps * 1e3
    return total


def tune_batch(specs, meshes, steps, ncurve=256, nsteps=300, use_cache=True, log=None):
    """Pick ``tile_k``, ``stream_nt`` and ``group_tetra`` for a population of cells by timing
    ``nsteps`` graph-replayed steps per candidate (the way the population then runs, launch costs
    included): one batch with each tile size and cache policy, and, when the population exceeds one
    cache-sized group, the cell groups of ``group_tetra()`` with each tile size (the default policy
    per group).  Returns ({"tile_k": K, "stream_nt": nt or None, "group_tetra": G or 0},
    {label: ms_per_step}), labels "256", "256nt", ..., "256g4" (4 groups)."""
    from nofluidx3d_amd.batch import BatchSimulation, CellGroups

    tetra = sum(len(m[1]) for m in meshes)
    key = f"batch4:{len(specs)}:{tetra}:{specs[0].viscosity_si is not None}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            return hit["best"], hit["ms"]
    times, choice = {}, {}
    for K in TILE_CANDIDATES:
        try:
            b = BatchSimulation(specs, meshes, steps=steps, ncurve=ncurve, tile_k=K)
        except Exception as e:  # a tile with more nodes than threads: candidate not applicable
            if log:
                log(f"tile_k {K}: {e}")
            continue
        for nt in STREAM_CANDIDATES:
            b.plan.stream_nt = nt
            lab = _label(K, nt)
            times[lab], choice[lab] = _time_steps(b, nsteps), {"tile_k": K, "stream_nt": nt, "group_tetra": 0}
            if log:
                log(f"{lab}: {times[lab]:.4f} ms/step")
        del b
    G = group_tetra()
    if tetra > G:
        for K in TILE_CANDIDATES:
            try:
                cg = CellGroups(specs, meshes, max_tetra=G, steps=steps, ncurve=ncurve, tile_k=K)
            except Exception as e:
                if log:
                    log(f"groups, tile_k {K}: {e}")
                continue
            lab = f"{K}g{len(cg.groups)}"
            times[lab], choice[lab] = _time_steps(cg, nsteps), {"tile_k": K, "stream_nt": None, "group_tetra": G}
            if log:
                log(f"{lab}: {times[lab]:.4f} ms/step")
            del cg
    best = choice[min(times, key=times.get)]
    _store(key, {"best": best, "ms": times, "when": time.time()})
    return best, times


def tune_simulation(sim, nsteps=300, log=None):
    """Pick ``tile_k``, ``stream_nt`` and the step form for a single-cell Simulation with the tiled
    scatter: each candidate runs ``nsteps`` steps to capture its graph, then ``nsteps`` graph-replayed
    steps timed on the host clock, launch costs included (what a run of the cell pays per step).
    The step forms are the node-fused one (one launch per step, when the plan admits it and
    ``sim.fuse_node_pass`` is not False) and two launches per step.  The state is restored afterwards.
    Sets ``sim.tile_k``, ``sim.stream_nt`` and ``sim.fuse_node_pass``; returns ((tile_k, stream_nt),
    {label: ms_per_step}) with labels "256", "256nt", ... and an "f" suffix for the node-fused form."""
    import torch

    from nofluidx3d_amd import _lib, openCL

    if sim.scatter != "tiled":
        return (sim.tile_k, sim.stream_nt), {}
    saved_time = sim.time
    sim.sync_forms()
    openCL.getCommandQueue().finish()
    snapshot = [t.clone() for t in sim.state_tensors()]  # timing runs real steps: restore afterwards
    forms = (False,) if sim.fuse_node_pass is False else (True, False)
    times, choice = {}, {}
    cands = ((128,) if sim.numTetra < SMALL_MESH_TETRA else ()) + TILE_CANDIDATES
    for K in cands:
        sim.tile_k = K
        sim._tiles = None
        for nt in STREAM_CANDIDATES:
            for fused in forms:
                sim.stream_nt, sim.fuse_node_pass = nt, fused
                sim._plan = None
                p = sim.plan()
                if p is None or sim.scatter != "tiled":
                    sim.scatter = "tiled"
                    break
                if fused and not p.disp_b:  # the plan cannot take the node-fused form
                    continue
                sim.advance(nsteps)
                _lib.call("nfx_stream_sync", openCL.handle())
                t0 = time.perf_counter()
                sim.advance(nsteps)
                _lib.call("nfx_stream_sync", openCL.handle())
                lab = _label(K, nt) + ("f" if fused else "")
                times[lab], choice[lab] = (time.perf_counter() - t0) / nsteps * 1e3, (K, nt, fused)
                if log:
                    log(f"{lab}: {times[lab]:.4f} ms/step")
    openCL.getCommandQueue().finish()
    for t, s0 in zip(sim.state_tensors(), snapshot):
        t.copy_(s0)
    torch.cuda.synchronize()
    sim.forms_restored()
    sim.time = saved_time
    K, nt, fused = choice[min(times, key=times.get)]
    sim.tile_k, sim.stream_nt, sim.fuse_node_pass = K, nt, fused
    sim._tiles = None
    sim._plan = None
    return (K, nt), times
