#!/usr/bin/env python3
"""Benchmark of the explicit tetrahedral-FEM step on one MI355X.

Workload (default): the large single cell -- R 10 um, lattice 0.2 um, seed 2 (~0.52M nodes,
~3.3M tetra), Mooney-Rivlin + viscoplastic, substrate and spherical tip, FP32 state.  A step is
one Simulation.timeStep of the whole cell (force pass + node pass, as two launches or, in the
node-fused form, as one launch carrying the previous step's node update: the autotuner picks the
form, the tile size and the cache policy); steps are replayed from a captured hipGraph.  Inputs are
resident in HBM before the timed region starts.

Printed JSON line (one):
- value = tetra-updates/s = T * steps / wall seconds of the timed region (the steps' graphs are
  captured during the warm-up, before the region; ``graph_captures_timed`` counts any inside it);
- hbm_gbs = algorithmic bytes per step / ms_per_step, with the per-unit bytes of the hot-path
  specification: 136 T + 108 N for the viscoplastic step (connectivity 16 B, Dm^-1 36 B, V0 4 B,
  the plastic state 36 B and epsbar 4 B each read and written, per tetra; x, v, f, f_old per node),
  56 T + 108 N for the elastic one;
- roofline: the dominant kernel (the force pass, ``k_tile_force``): those algorithmic bytes per
  launch (136 T + 24 N viscoplastic, 56 T + 24 N elastic; the whole step's bytes when the launch is
  node-fused) over its HIP-event average duration,
  against 8 TB/s; ``kernel_bytes_per_launch`` is what the kernel itself must move (it reads per-
  corner tile records instead of the connectivity and derives V0 from Dm^-1); ``traffic`` = HBM
  bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated) in a separate
  profiling run of this same command, read from ``--traffic-json`` (``traffic_source`` says which);
- cpu_baseline: the FP64 C oracle (OpenMP, all allowed host threads) on a bounded step sample
  of the same cell;
- other_step_form: the same window in the step form the tuner did not pick;
- plain_small: the small cell (R 8 um, 0.8 um lattice, ~25k tetra) fused step, for comparison;
- batched: the 512-cell workload (10k steps each, 256-point curve per cell);
- surrogate: MLP inference / training-step throughput on the MFMA kernels and a surrogate-assisted
  fit of 512 new cells confirmed by the plain simulator (when trained weights are present).

``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# FP32 vector peak of the VALUs without packed FP32 instructions (which this build does not emit):
# 256 CUs x 4 SIMDs x 16 lanes x 2 FLOP (FMA) x 2.4 GHz; AMD's 157.3 TFLOP/s figure counts packed FP32
VALU_FP32_PEAK_TFLOPS = 78.6
# algorithmic bytes per tetra of one force pass, as the hot-path specification counts them
SPEC_TET_BYTES = {"viscoplastic": 16 + 36 + 4 + 2 * 36 + 2 * 4, "elastic": 16 + 36 + 4}
# what the tiled force pass reads and writes per tetra: corner records 8 (instead of the 16 B
# connectivity), the plastic deviation Q 36 + 36 and epsbar 4 + 4 (viscoplastic), and, when it
# streams the reference shape instead of forming it from the reference positions, Dm^-1 36 and V0 4
# (elastic; the viscoplastic law takes V0 from det Dm^-1)
KERNEL_TET_BYTES = {"viscoplastic": 8 + 2 * 36 + 2 * 4, "elastic": 8}
STREAMED_SHAPE_BYTES = {"viscoplastic": 36, "elastic": 36 + 4}
# per node: u read 12, partial written 12, and the reference position X0 12 when the shape is formed
# from it (each node counted once; tile halos re-read some)
KERNEL_NODE_BYTES = 24


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cached_mesh(cfg):
    from nofluidx3d_amd import configs

    m = cfg["MESH"]
    path = f"/tmp/nfx_mesh_{m['radius_um']}_{m['spacing_um']}_{m['seed']}.npz"
    if os.path.exists(path):
        d = np.load(path)
        return d["P"], d["T"]
    t0 = time.time()
    P, T = configs.make_mesh(cfg)
    tmp = f"{path}.{os.getpid()}.tmp.npz"
    np.savez(tmp, P=P, T=T)
    os.replace(tmp, path)
    log(f"mesh {m}: {len(P)} nodes {len(T)} tetra in {time.time() - t0:.1f}s")
    return P, T


def make_sim(cfg, P, T, scatter, material, tile_k=512, shape_from_ref=True, compact_state=False, node_fused=None):
    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    sim = Simulation(cfg, scatter=scatter, mesh=vtk.from_tetra(P, T), verbose=False, tile_k=tile_k,
                     shape_from_ref=shape_from_ref, fuse_node_pass=node_fused)
    sim.setInteractionSubstrate()
    sim.setInteractionSphere()
    if material == "viscoplastic":
        sim.setInteractionFiniteStrainViscoplastic(compactState=compact_state)
    else:
        sim.setInteractionTetra()
    sim.setInteractionVelocityVerlet()
    return sim


def sync(sim):
    import torch

    from nofluidx3d_amd import _lib, openCL

    _lib.call("nfx_stream_sync", openCL.handle())
    torch.cuda.synchronize()


def graph_captures():
    import ctypes

    from nofluidx3d_amd import _lib, openCL

    n = ctypes.c_int64()
    _lib.call("nfx_graph_captures", openCL.handle(), ctypes.byref(n))
    return n.value


def time_steps(sim, warmup, steps):
    """ms per step over ``steps`` steps, after a first call of the same length (which captures the
    graph blocks the timed call replays) and ``warmup`` more steps; returns (s/step, captures in the
    timed region)"""
    sim.advance(steps)
    sim.advance(warmup)
    sync(sim)
    c0 = graph_captures()
    t0 = time.perf_counter()
    sim.advance(steps)
    sync(sim)
    dt = (time.perf_counter() - t0) / steps
    return dt, graph_captures() - c0


def pass_times(sim, nrep):
    import ctypes

    from nofluidx3d_amd import _lib, openCL

    mf, mn = ctypes.c_double(), ctypes.c_double()
    _lib.call("nfx_time_passes", openCL.handle(), sim.plan(), sim.time, nrep, ctypes.byref(mf), ctypes.byref(mn))
    sim.time += nrep
    sync(sim)
    return mf.value, mn.value


def cpu_baseline(P, T, cfg, material, target_s=12.0):
    import oracle as O
    from nofluidx3d_amd.TetraCell import TetraCell

    c = TetraCell.from_arrays(P, T, radius=cfg["CELL"]["RadiusSIM"])
    u = O.units(cfg["CELL"]["RadiusSI"], cfg["CELL"]["RadiusSIM"], cfg["CELL"]["YoungsModulusSI"], cfg["RhoSI"])
    prm = O.RunParams()
    E, nu = 0.01, cfg["CELL"]["PoissonRatio"]
    ratio = cfg["CELL"].get("MooneyRivlinRatio", 1.0)
    mu = E / (2 * (1 + nu))
    if material == "viscoplastic":
        prm.model = O.VISCOPLASTIC
        cc = cfg["CELL"]
        prm.vp = O.VPParams(ratio * mu, (1 - ratio) * mu, E / (3 * (1 - 2 * nu)), 1.0, 3 * cc["ViscositySI"] / u["p0"],
                            u["T0"], 1.0, 0.0, cc["YieldStressSI"] / u["p0"])
    else:
        prm.model, prm.E, prm.nu, prm.ratio = O.MR, E, nu, ratio
    prm.substrate, prm.substrate_k = 1, 0.1
    prm.wall = -(cfg["CELL"]["RadiusSIM"] + cfg.get("SubstrateDistance", cfg["InitialDistance"]))
    R = cfg["SphereRadiusSI"] / u["L0"]
    prm.indenter, prm.ind_k, prm.ind_radius = 1, 0.1, R
    prm.ind_start = cfg["CELL"]["RadiusSIM"] + R + cfg["InitialDistance"]
    prm.ind_speed, prm.ind_move_steps = cfg["VelocitySI"] / u["V0"], cfg["MoveTimeSI"] / u["T0"]
    threads = O.max_threads()
    prm.nthreads = threads
    st = O.OracleState(c.points, c.tetras, prm)
    t0 = time.perf_counter()
    st.run(1, record=False)
    t1 = time.perf_counter() - t0
    n = int(max(1, min(200, target_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    st.run(n, record=False)
    dt = time.perf_counter() - t0
    rate = len(T) * n / dt
    return {
        "value": rate,
        "unit": "tetra-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} FP64 steps of the same cell ({len(T)} tetra, {material}) with the C oracle, "
                  f"OpenMP {threads} threads, atomic scatter, {dt:.1f}s",
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "hbm_equiv_gbs": None,
    }


def bench_batched(steps=10000, ncells=512, form="auto", log=log):
    """The batched workload: 512 seeded cells (seeds 1000+i), 10k steps each, one 256-point curve per
    cell; tile size, cache policy and cell grouping from the autotuner (``form`` "auto"), or fixed:
    "groups" (cache-sized cell groups, tile 256, default policy) or "single" (one batch, tile 256,
    nontemporal).  Returns the JSON block."""
    from nofluidx3d_amd.autotune import group_tetra, tune_batch
    from nofluidx3d_amd.batch import CellGroups, make_meshes, sample_batch

    specs, stats = sample_batch(ncells, seed0=1000, steps=steps)
    t0 = time.time()
    meshes = make_meshes(specs)
    tmesh = time.time() - t0
    if form == "auto":
        best, tuned = tune_batch(specs, meshes, steps, use_cache=False, log=log)
    elif form == "groups":
        best, tuned = {"tile_k": 256, "stream_nt": None, "group_tetra": group_tetra()}, None
    else:
        best, tuned = {"tile_k": 256, "stream_nt": 1, "group_tetra": 0}, None
    b = CellGroups(specs, meshes, max_tetra=best.get("group_tetra") or None, steps=steps, ncurve=256,
                   tile_k=best["tile_k"], stream_nt=best["stream_nt"])
    b.advance(100)  # captures every group's 100-step graph; the groups keep them
    b.sync()
    c0 = graph_captures()
    t0 = time.perf_counter()
    b.advance(steps - 100)
    b.sync()
    dt = time.perf_counter() - t0
    captures_timed = graph_captures() - c0
    ms = dt / (steps - 100) * 1e3
    f, n, c = b.time_passes(20)
    curves = b.curves_normalized()
    ok = np.isfinite(curves).all(axis=1)
    tet = np.array([len(m[1]) for m in meshes])
    # algorithmic bytes as the specification counts a viscoplastic step (136 T + 108 N); the batch's
    # kernels move less: the linear Maxwell law does not read the accumulated plastic strain (not
    # kept), its plastic state is the symmetric deviation Gd (6 floats) instead of 9, and its node
    # state is (u, w) without the previous force
    step_bytes = SPEC_TET_BYTES["viscoplastic"] * b.T_real + 108 * b.N
    geo = b.shape_from_ref
    kernel_step_bytes = (8 + (0 if geo else 36) + 2 * (24 if b.sym_state else 36) + (8 if b.plastic_strain else 0)) * \
        b.T_real + (2 * 24 + 4 + 4 + 12 + (12 if geo else 0)) * b.N
    # per node: u and w read and written, X0.y, node_cell; u (and X0 when the shape is formed from it)
    # read by the tiles
    return {
        "workload": f"{ncells} cells (seeds 1000+i), {b.T_real} tetra, {b.N} nodes, NH + Maxwell viscoplastic, "
                    f"{steps} steps each, 256-point curve per cell",
        "cells": ncells,
        "tetra": int(b.T_real),
        "nodes": int(b.N),
        "tetra_per_cell": {"min": int(tet.min()), "median": int(np.median(tet)), "max": int(tet.max())},
        "sampling_redraws": stats,
        "tile_k": best["tile_k"],
        "stream_nt": b.stream_nt if len(b.groups) > 1 else b.stream_nt[0],
        "cell_groups": len(b.groups),
        "group_tetra": int(best.get("group_tetra") or 0),
        "autotune_ms_per_step": tuned,
        "algorithmic_bytes_per_step": int(step_bytes),
        "kernel_bytes_per_step": int(kernel_step_bytes),
        "ms_per_step": ms,
        "graph_captures_timed": captures_timed,
        "value": b.T_real * 1e3 / ms,
        "unit": "tetra-updates/s",
        "hbm_gbs": step_bytes / (ms * 1e-3) / 1e9,
        "ms_force_pass": f,
        "ms_node_pass": n,
        "ms_curve_pass": c,
        "full_run_s": dt * steps / (steps - 100),
        "cell_curves_per_s": ncells / (dt * steps / (steps - 100)),
        "cells_finite": int(ok.sum()),
        "mesh_generation_s": tmesh,
    }


def bench_surrogate(weights, n_targets=4096, n_nm=16, nm_evals=300, single=16, single_nm=1, log=log):
    """The learned part: MLP inference and training-step throughput on the MFMA kernels, then the
    fitting benchmark (``surrogate.fit_bench``): ``n_targets`` new cells (batched workload generator,
    seeds 7000+i, 1% noise) fitted together on the surrogate path with the plain simulator refining
    and confirming, simulation-only Nelder-Mead on the first ``n_nm`` of them for ``nm_evals``
    evaluations (the evaluations it needs for the surrogate path's median error are reported), and
    the p50/p99 latency of one-target fits."""
    import torch

    from nofluidx3d_amd.surrogate import fit_bench, train

    net, norm = train.load(weights)
    sizes = net.sizes
    flops_row = 2 * sum(a * b for a, b in zip(net.in_pad, sizes[1:]))
    rows = 65536
    X = np.random.default_rng(0).normal(size=(rows, sizes[0])).astype(np.float32)
    net.set_input(X)
    net.forward_dev(rows)
    net.sync()
    t0 = time.perf_counter()
    for _ in range(10):
        net.forward_dev(rows)
    net.sync()
    t_inf = (time.perf_counter() - t0) / 10
    B = 4096
    net.set_input(X[:B], train=True)
    tgt = torch.zeros(B, sizes[-1], dtype=torch.float32, device="cuda:0")
    snap = net.state()
    net.train_step(B, tgt)
    net.sync()
    t0 = time.perf_counter()
    for _ in range(20):
        net.train_step(B, tgt)
    net.sync()
    t_train = (time.perf_counter() - t0) / 20
    net.load_state(snap)
    out = {"weights": os.path.relpath(weights, ROOT), "sizes": list(sizes),
           "inference_rows_per_s": rows / t_inf, "inference_tflops": flops_row * rows / t_inf / 1e12,
           "train_step_ms_batch4096": t_train * 1e3, "train_tflops": 3 * flops_row * B / t_train / 1e12,
           "mfma_peak_tflops_bf16_dense": 2500.0}
    args = fit_bench.options(["--weights", weights, "--n", str(n_targets), "--n-nm", str(n_nm), "--nm-evals",
                              str(nm_evals), "--single", str(single), "--single-nm", str(single_nm)])
    out["fit"] = fit_bench.run(net, norm, args, log=log)
    return out


def stream_copy_gbs(nbytes=1 << 30, nrep=20):
    """Measured copy bandwidth (read + write bytes / time) with the library's vectorized copy."""
    import ctypes

    import torch

    from nofluidx3d_amd import _lib, openCL

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    ms = ctypes.c_double()
    _lib.call("nfx_stream_copy", openCL.handle(), a.data_ptr(), b.data_ptr(), nbytes, nrep, ctypes.byref(ms))
    ok = bool(torch.equal(a[:1024], b[:1024]))
    del a, b
    return {"gbs": 2 * nbytes / (ms.value * 1e-3) / 1e9, "bytes_copied": nbytes, "ms": ms.value, "verified": ok}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="large", choices=["large", "small"])
    ap.add_argument("--material", default=None, choices=[None, "viscoplastic", "elastic"])
    ap.add_argument("--scatter", default="tiled", choices=["tiled", "csr", "atomic"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched 512-cell measurement")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate measurement")
    ap.add_argument("--tile-k", type=int, default=0, choices=[0, 128, 256, 512], help="0: autotune")
    ap.add_argument("--start-step", type=int, default=-1,
                    help="protocol step the timed window starts from (untimed steps before it); -1: the end of "
                         "the indentation for the large cell (the deepest contact, the largest plastic zone), 0 "
                         "for the small cell")
    ap.add_argument("--no-full-protocol", action="store_true", help="skip timing the whole protocol")
    ap.add_argument("--save-state", default=None, metavar="NPZ",
                    help="advance to the start step, save the simulation state there and exit")
    ap.add_argument("--load-state", default=None, metavar="NPZ",
                    help="start the timed window from a state --save-state wrote (same workload)")
    ap.add_argument("--compact-state", action="store_true",
                    help="viscoplastic: the fused tiled step keeps the symmetric plastic deviation only "
                         "(same forces; no Fp)")
    ap.add_argument("--step-form", choices=("auto", "fused", "two"), default="auto",
                    help="fused: the node update of each step rides in the next force pass (one launch per "
                         "step); two: a force pass and a node pass; auto: the autotuner's choice, or by mesh "
                         "size with a fixed --tile-k")
    ap.add_argument("--streamed-shape", action="store_true",
                    help="the tiled force pass reads the precomputed Dm^-1 and V0 instead of forming them "
                         "from the reference positions")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic_large_cell.json"))
    ap.add_argument("--valu-json", default=os.path.join(ROOT, "profiles", "valu_large_cell.json"))
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="also run one device-to-device copy of BYTES (HBM-counter calibration under rocprofv3)")
    ap.add_argument("--only", default=None, choices=[None, "batched", "surrogate"],
                    help="run just one secondary measurement and print its block (profiling runs)")
    ap.add_argument("--batched-steps", type=int, default=10000)
    ap.add_argument("--batched-form", default="auto", choices=["auto", "groups", "single"],
                    help="batched population as tuned, as cache-sized cell groups, or as one batch (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is design-only "
              f"in this build.", file=sys.stderr)
        sys.exit(2)

    import torch

    from nofluidx3d_amd import configs

    assert torch.cuda.is_available(), "no HIP device"
    if args.calibrate_copy:
        a = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.only == "batched":
        print(json.dumps(bench_batched(steps=args.batched_steps, form=args.batched_form)), flush=True)
        return
    if args.only == "surrogate":
        print(json.dumps(bench_surrogate(os.path.join(ROOT, "nofluidx3d_amd", "surrogate", "weights",
                                                      "surrogate.npz"))), flush=True)
        return
    if args.workload == "large":
        cfg = configs.large_cell()
        material = args.material or "viscoplastic"
    else:
        cfg = configs.small_cell()
        material = args.material or "elastic"
    P, T = cached_mesh(cfg)
    N, nT = len(P), len(T)
    compact = args.compact_state and material == "viscoplastic" and args.scatter == "tiled"
    sim = make_sim(cfg, P, T, args.scatter, material, args.tile_k or 512, not args.streamed_shape, compact,
                   {"auto": None, "fused": True, "two": False}[args.step_form])
    assert sim.plan() is not None
    move_steps = int(round(cfg["MoveTimeSI"] / sim.T0))
    start = (move_steps if args.workload == "large" else 0) if args.start_step < 0 else args.start_step
    if args.load_state:  # the protocol state a previous run saved (a profiler then sees the window only)
        sim.load_state(args.load_state)
        start = sim.time
        log(f"loaded step {start} from {args.load_state}")
    elif start > 0:  # untimed: bring the cell to the protocol state the timed window starts from
        t0 = time.perf_counter()
        sim.advance(start)
        sync(sim)
        log(f"advanced to step {start} in {time.perf_counter() - t0:.1f}s")
    if args.save_state:
        sim.save_state(args.save_state)
        print(json.dumps({"saved_state": args.save_state, "step": sim.time}), flush=True)
        return
    tuned = None
    if args.scatter == "tiled" and args.tile_k == 0:  # tuned in the state the window starts from
        from nofluidx3d_amd.autotune import tune_simulation

        _, tuned = tune_simulation(sim, log=log)
    dt, captures_timed = time_steps(sim, args.warmup, args.steps)
    ms = dt * 1e3
    ms_force, ms_node = pass_times(sim, min(args.steps, 100))
    vp = material == "viscoplastic"
    plastic_tiles = sim.plastic_tile_fraction()
    # node-fused: the dominant kernel is the force pass carrying the previous step's node update, so
    # one launch moves a whole step's algorithmic bytes
    node_fused = bool(sim.plan().disp_b)
    step_bytes = SPEC_TET_BYTES[material] * nT + 108 * N
    force_bytes = step_bytes if node_fused else SPEC_TET_BYTES[material] * nT + 24 * N
    geo = bool(sim.plan().shape_from_ref) and args.scatter == "tiled"
    tet_bytes = KERNEL_TET_BYTES[material] + (0 if geo else STREAMED_SHAPE_BYTES[material])
    if compact:  # Gd 24 + 24 instead of Q 36 + 36; epsbar only when the flow reads it
        tet_bytes -= 2 * 12 + (0 if sim.plan().epsbar else 8)
    kernel_bytes = tet_bytes * nT + (KERNEL_NODE_BYTES + (12 if geo else 0)) * N
    if node_fused:  # the node update: the previous partials (12 B each, M of them), node_off 4, w read,
        # u and w written (12 each) per node
        kernel_bytes += 12 * sim.node_tiles()["m_total"] + (4 + 3 * 12) * N
    achieved = force_bytes / (ms_force * 1e-3) / 1e9
    nt_flag = sim.plan().stream_nt
    kernel = (f"k_tile_force<{'NFX_VISCOPLASTIC' if vp else 'NFX_MR'}, K={sim.tile_k}, "
              f"{'shape from reference positions' if geo else 'streamed shape'}, "
              f"{'nontemporal' if nt_flag else 'default'} cache policy"
              f"{', node-fused (carries the previous step node update)' if node_fused else ''}>") \
        if args.scatter == "tiled" else f"{'k_vp' if vp else 'k_tetra'}<{args.scatter}>"
    prof_key = (f"{material}_{args.scatter}_{sim.tile_k}{'nt' if nt_flag else ''}{'_shape_from_ref' if geo else ''}"
                f"{'_node_fused' if node_fused else ''}")
    traffic, traffic_source = None, None
    if args.traffic_json and os.path.exists(args.traffic_json):
        tj = json.load(open(args.traffic_json))
        key = prof_key
        if tj.get("tetra") == nT and key in tj.get("force_pass_bytes", {}) and start == tj.get("from_step"):
            traffic = tj["force_pass_bytes"][key]
            traffic_source = f"{os.path.relpath(args.traffic_json, ROOT)}[{key}]: {tj.get('source', '')}"
    compute = None
    if args.valu_json and os.path.exists(args.valu_json):
        vj = json.load(open(args.valu_json))
        key = prof_key
        if vj.get("tetra") == nT and key in vj.get("force_pass", {}) and start == vj.get("from_step"):
            fl = vj["force_pass"][key]["fp32_flops"]
            compute = {"bound": "valu", "flops_per_launch": fl, "achieved": fl / (ms_force * 1e-3) / 1e12,
                       "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": fl / (ms_force * 1e-3) / 1e12 / VALU_FP32_PEAK_TFLOPS,
                       "valu_insts_per_wave": vj["force_pass"][key].get("valu_insts_per_wave"),
                       "source": f"{os.path.relpath(args.valu_json, ROOT)}[{key}]: {vj.get('source', '')}"}
    out = {
        "metric": "tetra-updates/s",
        "value": nT * 1e3 / ms,
        "unit": "tetra-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "graph_captures_timed": captures_timed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}_cell: R {cfg['MESH']['radius_um']} um, lattice {cfg['MESH']['spacing_um']} "
                        f"um, seed {cfg['MESH']['seed']}, {N} nodes, {nT} tetra, "
                        f"{'Mooney-Rivlin + viscoplastic' if vp else cfg['CELL']['model']}, substrate + sphere tip, "
                        f"fused step, {args.scatter} scatter, hipGraph replay, timed from protocol step {start}"
                        + (" (end of the 3 um indentation)" if start == move_steps and args.workload == "large" else ""),
            "nodes": N,
            "tetra": nT,
            "scatter": args.scatter,
            "tile_k": sim.tile_k,
            "stream_nt": sim.stream_nt,
            "timed_from_step": start,
            "protocol_steps": int(sim.Nsteps),
            "shape_from_ref": geo,
            "step_form": "node-fused (one launch per step)" if node_fused else "two launches per step",
            "viscoplastic_state": ("compact: symmetric deviation Gd" + ("" if sim.plan().epsbar else ", no plastic strain")
                                   if compact else "Q = Fp^-1 - I and plastic strain") if vp else None,
            "autotune_ms_per_step": tuned,
            "parallelism": "single GPU",
        },
        "hbm_gbs": step_bytes / (ms * 1e-3) / 1e9,
        "hbm_frac": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "ms_force_pass": ms_force,
        "ms_node_pass": ms_node,
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "algorithmic_bytes_per_launch": force_bytes,
            "node_fused": node_fused,
            "kernel_bytes_per_launch": kernel_bytes,
            "plastic_tile_fraction": plastic_tiles,
            "note": (f"achieved counts the specification's bytes per launch (136 B/tetra + "
                     f"{108 if node_fused else 24} B/node{': the launch carries the previous step node update, so a step is one launch' if node_fused else ''}); the kernel "
                     "forms Dm^-1 from the reference positions and touches the plastic state only in tiles that "
                     "have yielded and tetra that flow, so it moves far less (traffic, measured) and is bound by "
                     "VALU issue (profiles/instructions_large_cell_viscoplastic_tiled.txt)") if vp and geo else None,
            "ms_per_launch": ms_force,
            "compute": compute,
        },
    }
    if args.scatter == "tiled" and not args.no_extra:  # the same cell and window in the other step form
        sim.fuse_node_pass = not node_fused
        if bool(sim.plan().disp_b) != node_fused:
            ms6 = time_steps(sim, args.warmup, min(args.steps, 1000))[0] * 1e3
            f6, n6 = pass_times(sim, 50)
            out["other_step_form"] = {"step": "two launches per step" if node_fused else
                                      "node-fused (one launch per step)", "ms_per_step": ms6,
                                      "value": nT * 1e3 / ms6, "ms_force_pass": f6, "ms_node_pass": n6}
        sim.fuse_node_pass = node_fused
    sim_tile_k, sim_stream_nt = sim.tile_k, sim.stream_nt
    del sim
    torch.cuda.empty_cache()
    if not args.no_extra:
        out["other_scatter"] = []
        for other in ("tiled", "csr", "atomic"):
            if other == args.scatter:
                continue
            s2 = make_sim(cfg, P, T, other, material)
            ms2 = time_steps(s2, args.warmup, min(args.steps, 200))[0] * 1e3
            f2, n2 = pass_times(s2, 50)
            out["other_scatter"].append({"scatter": other, "ms_per_step": ms2, "value": nT * 1e3 / ms2,
                                         "ms_force_pass": f2, "ms_node_pass": n2})
            del s2
        if vp and args.scatter == "tiled":  # the other viscoplastic state form, same tile shape and window
            s4 = make_sim(cfg, P, T, "tiled", material, sim_tile_k, True, not compact, node_fused)
            s4.stream_nt = sim_stream_nt
            if start > 0:
                s4.advance(start)
            ms4 = time_steps(s4, args.warmup, min(args.steps, 200))[0] * 1e3
            f4, n4 = pass_times(s4, 50)
            out["other_state_form"] = {"viscoplastic_state": "Q = Fp^-1 - I and plastic strain" if compact else
                                       "compact: symmetric deviation Gd", "ms_per_step": ms4,
                                       "value": nT * 1e3 / ms4, "ms_force_pass": f4, "ms_node_pass": n4}
            del s4
        small = configs.small_cell()
        Ps, Ts = cached_mesh(small)
        s3 = make_sim(small, Ps, Ts, "tiled", "elastic")
        from nofluidx3d_amd.autotune import tune_simulation

        (k3, nt3), tuned3 = tune_simulation(s3)
        ms3 = time_steps(s3, 200, 2000)[0] * 1e3
        f3, n3 = pass_times(s3, 200)
        out["plain_small"] = {"workload": f"small_cell {len(Ps)} nodes {len(Ts)} tetra NH", "ms_per_step": ms3,
                              "value": len(Ts) * 1e3 / ms3, "unit": "tetra-updates/s", "tile_k": k3,
                              "node_fused": bool(s3.plan().disp_b),
                              "ms_force_pass": f3, "ms_node_pass": n3, "autotune_ms_per_step": tuned3}
        del s3
    if not args.no_extra and not args.no_full_protocol and args.scatter == "tiled":
        # the whole protocol from rest on a fresh cell, graph captures included (one per distinct
        # block length, milliseconds against seconds)
        s5 = make_sim(cfg, P, T, "tiled", material, sim_tile_k, not args.streamed_shape, compact, node_fused)
        s5.stream_nt = sim_stream_nt
        n5 = int(s5.Nsteps)
        sync(s5)
        t0 = time.perf_counter()
        s5.advance(n5)
        sync(s5)
        dt5 = time.perf_counter() - t0
        out["full_protocol"] = {"steps": n5, "seconds": dt5, "ms_per_step": dt5 / n5 * 1e3,
                                "value": nT * n5 / dt5, "unit": "tetra-updates/s",
                                "finite": bool(np.isfinite(s5.positions()).all())}
        del s5
    if not args.no_extra:
        sc = stream_copy_gbs()
        out["stream_copy"] = sc
        out["roofline"]["frac_of_stream_copy"] = achieved / sc["gbs"]
    if not args.no_batched and not args.no_extra:
        out["batched"] = bench_batched(steps=args.batched_steps)
        if "plain_small" in out:
            out["batched"]["x_plain_small"] = out["batched"]["value"] / out["plain_small"]["value"]
    weights = os.path.join(ROOT, "nofluidx3d_amd", "surrogate", "weights", "surrogate.npz")
    if not args.no_extra and not args.no_surrogate and os.path.exists(weights):
        out["surrogate"] = bench_surrogate(weights)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, T, cfg, material)
        out["cpu_baseline"]["hbm_equiv_gbs"] = out["cpu_baseline"]["value"] / nT * step_bytes / 1e9
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
