"""Batched multi-cell AFM indentation: many cells with their own meshes and parameters advance
together on one GPU.

The reference simulates one cell per process (``Simulation.py``); this module runs a population
of cells as one concatenated mesh.  Tiles of the tiled scatter are aligned to cells, so every cell
sums its nodal forces in exactly the order it would use alone, and a cell's batched trajectory
equals its single-cell run up to rounding (the batch carries the damped velocity-Verlet state as
(x, u = a v + f/2), the same recurrence without the previous force).  Per step: one force pass over
all tiles, one node pass over all nodes and a curve pass (one workgroup per cell: on sampling
steps the cell's FP64 indenter force, otherwise an early exit); steps replay from a captured
hipGraph (``nfx_batch_run_steps``).

Cell model (per cell, in that cell's simulation units, ``Simulation.py:42-56``): Neo-Hookean
elasticity with E_sim = 0.01 and the cell's Poisson ratio, optionally the Maxwell-type
viscoplastic flow (m = 1, n = 0, Y = 3 eta * 1/s, rate = T0 per step; ``yield_si`` offsets the
von Mises stress); a substrate wall touching the cell's bottom; a spherical tip touching its top at
step 0 and moving down at the cell's speed for the whole run, to ``indent_frac * R``.  The lattice
spacing sets RadiusSIM = (R / spacing) * h with h <= 1 simulation lengths per lattice spacing,
small enough for explicit stability (``CellSpec.lattice_sim``); mass scaling sets
T0 = (indentation time) / steps.  The curve is the tip force sampled ``ncurve`` times, normalised by
E R^2 (F_sim / (0.01 RadiusSIM^2)).
"""

from __future__ import annotations

import ctypes
import dataclasses
import math
import os

import numpy as np

from nofluidx3d_amd import _lib
from nofluidx3d_amd.autotune import default_stream_nt

E_SIM = 1e-2


@dataclasses.dataclass
class CellSpec:
    radius_um: float
    spacing_um: float
    seed: int
    E_si: float
    nu: float
    viscosity_si: float | None  # None: purely elastic
    tip_radius_um: float
    speed_um_s: float
    yield_si: float = 0.0
    indent_frac: float = 0.3

    def lattice_sim(self):
        """Lattice spacing in simulation units.  The explicit step (dt = 1, unit node mass) is stable
        while the largest stiffness eigenvalue stays below 4; for the seeded meshes (quality >= 0.1)
        it is about 27 * M * h with M = E (1 - nu) / ((1 + nu)(1 - 2 nu)) the P-wave modulus in
        simulation units, so h = min(1, 0.075 / M) keeps it at or below 2."""
        M = E_SIM * (1 - self.nu) / ((1 + self.nu) * (1 - 2 * self.nu))
        return min(1.0, 0.075 / M)

    def units(self, steps):
        """Simulation units of this cell: RadiusSIM, L0, p0, T0, V0."""
        m = max(int(round(self.radius_um / self.spacing_um)), 2)
        radius_sim = float(m) * self.lattice_sim()
        L0 = self.radius_um * 1e-6 / radius_sim
        p0 = self.E_si / E_SIM
        t_ind = self.indent_frac * self.radius_um / self.speed_um_s
        T0 = t_ind / steps
        V0 = L0 / T0
        return {"RadiusSIM": radius_sim, "L0": L0, "p0": p0, "T0": T0, "V0": V0, "rho": p0 / V0**2}

    def relaxation_per_step(self, steps):
        """G * T0 / eta: the fraction of deviatoric stress the explicit Maxwell update removes per step."""
        if self.viscosity_si is None:
            return 0.0
        G = self.E_si / (2 * (1 + self.nu))
        return G * self.units(steps)["T0"] / self.viscosity_si


def surrogate_inputs(E_si, nu, viscosity_si, tip_radius_um, speed_um_s, radius_um, spacing_um):
    """The six surrogate inputs (n, 6): log10 (viscosity / E) [s], nu, log10 tip radius [um],
    log10 speed [um/s], log10 cell radius [um], log10 lattice spacing [um].
    - E and the viscosity enter only through their ratio: scaling both by one factor leaves the
      run in simulation units unchanged, so the curve normalised by E R^2 is invariant.
    - The lattice spacing sets the mesh resolution, which moves the curve by several per cent at
      fixed physics.
    In logs, the dimensionless groups of the problem (tip / cell radius, the Deborah number
    viscosity * speed / (E * radius), radius / spacing) are linear combinations of the inputs."""
    f = lambda v: np.asarray(v, np.float64)  # noqa: E731
    return np.stack([np.log10(f(viscosity_si) / f(E_si)), f(nu), np.log10(f(tip_radius_um)), np.log10(f(speed_um_s)),
                     np.log10(f(radius_um)), np.log10(f(spacing_um))], 1)


def estimated_tetra(radius_um, spacing_um):
    """~6 tetra per lattice node for the seeded generator (quality filter included)."""
    m = max(int(round(radius_um / spacing_um)), 2)
    return 6.05 * 4.0 / 3.0 * math.pi * m**3


def sample_batch(n=512, seed0=1000, steps=10000, t_range=(2000, 60000), max_relax=0.1, viscous=True):
    """Per-cell seeds seed0 + i.  R ~ U(4, 10) um, spacing ~ U(0.3, 0.6) um, E log-uniform 0.1-10 kPa,
    viscosity log-uniform 1-1000 Pa s, nu ~ U(0.3, 0.49), tip radius ~ U(1, 5) um, speed log-uniform
    0.5-50 um/s.  (R, spacing) is redrawn until the cell has ``t_range`` tetra (the stated 2k-60k),
    and (E, nu, viscosity, speed) until the explicit Maxwell update removes at most ``max_relax`` of
    the deviatoric stress per step (stability of the reference's one-step explicit flow update)."""
    specs = []
    stats = {"redraw_size": 0, "redraw_relax": 0}
    for i in range(n):
        rng = np.random.Generator(np.random.PCG64(seed0 + i))
        while True:
            R, s = rng.uniform(4.0, 10.0), rng.uniform(0.3, 0.6)
            if t_range[0] <= estimated_tetra(R, s) <= t_range[1]:
                break
            stats["redraw_size"] += 1
        tip = rng.uniform(1.0, 5.0)
        while True:
            E = 10 ** rng.uniform(2.0, 4.0)
            eta = 10 ** rng.uniform(0.0, 3.0) if viscous else None
            nu = rng.uniform(0.3, 0.49)
            v = 10 ** rng.uniform(math.log10(0.5), math.log10(50.0))
            spec = CellSpec(R, s, seed0 + i, E, nu, eta, tip, v)
            if spec.relaxation_per_step(steps) <= max_relax:
                break
            stats["redraw_relax"] += 1
        specs.append(spec)
    return specs, stats


def _mesh_path(R, s, seed, cache):
    return os.path.join(cache, f"cell_{R:.6f}_{s:.6f}_{seed}.npz") if cache else None


def _mesh_job(args):
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    R, s, seed, cache = args
    path = _mesh_path(R, s, seed, cache)
    if path and os.path.exists(path):
        d = np.load(path)
        return d["P"], d["T"]
    P, T = generate_sphere_mesh(R, s, seed)
    if path:  # write then rename: an interrupted run leaves no truncated cache file
        tmp = f"{path}.{os.getpid()}.tmp.npz"
        np.savez(tmp, P=P, T=T)
        os.replace(tmp, path)
    return P, T


def make_meshes(specs, workers=None, cache="/tmp/nfx_cells"):
    """Seeded meshes of all cells.  Cached meshes are read in this process; only the missing ones
    are generated, by a spawned process pool (no GPU use in the workers), so a run whose meshes are
    all cached starts no worker (and a profiler preloaded into this process reaches none)."""
    import multiprocessing as mp

    if cache:
        os.makedirs(cache, exist_ok=True)
    jobs = [(c.radius_um, c.spacing_um, c.seed, cache) for c in specs]
    missing = [i for i, j in enumerate(jobs) if not (cache and os.path.exists(_mesh_path(*j)))]
    out = [None] * len(jobs)
    absent = set(missing)
    for i, j in enumerate(jobs):
        if i not in absent:
            out[i] = _mesh_job(j)
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(missing) < 4:
        for i in missing:
            out[i] = _mesh_job(jobs[i])
        return out
    # close + join: the workers end on their own rather than by SIGTERM (a profiler preloaded into
    # them handles that signal and can stall the teardown)
    pool = mp.get_context("spawn").Pool(workers)
    try:
        made = pool.map(_mesh_job, [jobs[i] for i in missing], chunksize=4)
        pool.close()
        pool.join()
    finally:
        pool.terminate()
    for i, m in zip(missing, made):
        out[i] = m
    return out


def sim_points(P_um, radius_sim):
    """TetraCell's recentre and scale (TetraCell.py:15-20) in FP64, held at the FP32 values the
    device keeps as the reference positions (as ``TetraCell`` holds them)."""
    p = np.asarray(P_um, np.float64) - P_um[0]
    return (p * (radius_sim / abs(p[1, 0]))).astype(np.float32).astype(np.float64)


def narrow_tile_arrays(tile_nodes, tile_dst, tile_count, MX):
    """The tile-node ids and partial positions as 16-bit offsets from per-tile bases (the first
    tile-node's: ids ascend within a tile, and so do the node-ordered partial positions).  A tile
    never leaves its cell, so the offsets fit when every cell has fewer than 65536 nodes; returns
    {} (the force pass keeps the 32-bit arrays) when one does not."""
    nodes = tile_nodes.reshape(-1, MX).astype(np.int64)
    dst = tile_dst.reshape(-1, MX).astype(np.int64)
    real = np.arange(MX)[None, :] < tile_count[:, None]
    nb, db = nodes[:, 0], dst[:, 0]
    on = np.where(real, nodes - nb[:, None], 0)
    od = np.where(real, dst - db[:, None], 0)
    if on.min() < 0 or od.min() < 0 or on.max() > 0xFFFF or od.max() > 0xFFFF:
        return {}
    return {"tile_nodes16": on.astype(np.uint16).reshape(-1), "tile_dst16": od.astype(np.uint16).reshape(-1),
            "tile_node_base": nb.astype(np.uint32), "tile_dst_base": db.astype(np.uint32)}


def fill_cell(c, spec, u, steps):
    """One nfx_cell record: the material (E_sim = 0.01, Neo-Hookean, optional Maxwell flow), the
    substrate and the tip law in the cell's simulation units ``u`` (the constants the single-cell
    path computes)."""
    Rs = u["RadiusSIM"]
    mu = E_SIM / (2 * (1 + spec.nu))
    c["mu1"], c["mu2"], c["kappa"] = mu, 0.0, E_SIM / (3 * (1 - 2 * spec.nu))
    if spec.viscosity_si is not None:
        c["vp_rate"] = u["T0"]
        c["vp_yield"] = 3.0 * spec.viscosity_si * 1.0 / u["p0"]
        c["vp_offset"] = spec.yield_si / u["p0"]
    tipR = spec.tip_radius_um * 1e-6 / u["L0"]
    c["wall"], c["substrate_k"] = -Rs, 0.1
    c["tip_radius"], c["tip_k"] = tipR, 0.1
    c["tip_start"] = Rs + tipR
    c["tip_speed"] = spec.speed_um_s * 1e-6 / u["V0"]
    c["tip_move_steps"] = float(steps)


def cell_table(specs, steps):
    """The nfx_cell records, units and RadiusSIM of ``specs`` (no geometry)."""
    cells = np.zeros(len(specs), dtype=CELL_DTYPE)
    units = [s.units(steps) for s in specs]
    for i, (spec, u) in enumerate(zip(specs, units)):
        fill_cell(cells[i], spec, u, steps)
    return cells, units, np.array([u["RadiusSIM"] for u in units])


def build_host_arrays(specs, meshes, steps, tile_k):
    """All host-side arrays of a batch (numpy), before upload.  Tiles are built per cell with a
    common capacity MX and concatenated with node/partial offsets."""
    K = int(tile_k)
    nc = len(specs)
    cells = np.zeros(nc, dtype=CELL_DTYPE)
    pts, evs, vols, tets, tile_meta, units = [], [], [], [], [], []
    node_off, tet_pad_off = [0], [0]
    radius_sim = np.zeros(nc)
    for i, (spec, (P, T)) in enumerate(zip(specs, meshes)):
        u = spec.units(steps)
        units.append(u)
        Rs = u["RadiusSIM"]
        radius_sim[i] = Rs
        x = sim_points(P, Rs)
        p = x[T]
        R1, R2, R3 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]
        ev = np.stack([R1, R2, R3], 1).reshape(len(T), 9)
        vol = np.einsum("ij,ij->i", R1, np.cross(R2, R3)) / 6
        nT = len(T)
        nTp = -(-nT // K) * K
        pad = nTp - nT
        # padding tetra (never read by the kernels): identity shape, the cell's first node
        evs.append(np.concatenate([ev, np.tile(np.eye(3).reshape(1, 9), (pad, 1))]))
        vols.append(np.concatenate([vol, np.full(pad, 1.0 / 6.0)]))
        tets.append(np.concatenate([T + node_off[-1], np.full((pad, 4), node_off[-1])]))
        pts.append(x)
        tile_meta.append((T, len(P), nT))
        node_off.append(node_off[-1] + len(P))
        tet_pad_off.append(tet_pad_off[-1] + nTp)
        fill_cell(cells[i], spec, u, steps)
    N = int(node_off[-1])

    vp = ctypes.c_void_p
    per = []
    MX = 0
    for T, n, nT in tile_meta:
        tt = np.ascontiguousarray(T.reshape(-1, order="F").astype(np.uint32))
        ntiles = -(-nT // K)
        cnt = np.zeros(ntiles, np.uint32)
        m, mx = ctypes.c_int64(), ctypes.c_int32()
        _lib.call("nfx_build_tiles_host", tt.ctypes.data_as(vp), n, nT, K, ctypes.byref(mx), ctypes.byref(m),
                  cnt.ctypes.data_as(vp), None, None, None, None, None)
        per.append((tt, n, nT, ntiles, cnt, m.value))
        MX = max(MX, mx.value)
    if MX > min(K, _lib.TILE_MAX_NODES):
        raise _lib.NfxError(f"a tile touches {MX} nodes > min(tile_k {K}, {_lib.TILE_MAX_NODES}); use another tile_k")
    nt_tot = sum(p[3] for p in per)
    h = {
        "tile_count": np.zeros(nt_tot, np.uint32),
        "tile_nodes": np.zeros(nt_tot * MX, np.uint32),
        "tile_lend": np.zeros(nt_tot * MX, np.uint16),
        "tile_corners": np.zeros(4 * K * nt_tot, np.uint16),
        "tile_dst": np.zeros(nt_tot * MX, np.uint32),
        "tile_ntet": np.zeros(nt_tot, np.uint16),
        "tile_cell": np.zeros(nt_tot, np.uint32),
        "node_poff": np.zeros(N + 1, np.uint32),
    }
    tpos, ppos = 0, 0
    for i, (tt, n, nT, ntiles, cnt, M) in enumerate(per):
        tn = np.zeros(ntiles * MX, np.uint32)
        tl = np.zeros(ntiles * MX, np.uint16)
        ts = np.zeros(4 * K * ntiles, np.uint16)
        td = np.zeros(ntiles * MX, np.uint32)
        no = np.zeros(n + 1, np.uint32)
        m, mx = ctypes.c_int64(M), ctypes.c_int32(MX)
        _lib.call("nfx_build_tiles_host", tt.ctypes.data_as(vp), n, nT, K, ctypes.byref(mx), ctypes.byref(m),
                  cnt.ctypes.data_as(vp), tn.ctypes.data_as(vp), tl.ctypes.data_as(vp), ts.ctypes.data_as(vp),
                  td.ctypes.data_as(vp), no.ctypes.data_as(vp))
        sl = slice(tpos, tpos + ntiles)
        h["tile_count"][sl] = cnt
        h["tile_nodes"][tpos * MX:(tpos + ntiles) * MX] = tn + node_off[i]
        h["tile_lend"][tpos * MX:(tpos + ntiles) * MX] = tl
        h["tile_corners"][4 * K * tpos:4 * K * (tpos + ntiles)] = ts
        h["tile_dst"][tpos * MX:(tpos + ntiles) * MX] = td + ppos
        h["tile_ntet"][sl] = np.minimum(K, nT - K * np.arange(ntiles))
        h["tile_cell"][sl] = i
        h["node_poff"][node_off[i]:node_off[i] + n + 1] = no + ppos
        tpos += ntiles
        ppos += M
    h.update(narrow_tile_arrays(h["tile_nodes"], h["tile_dst"], h["tile_count"], MX))
    h.update({
        "cells": cells, "units": units, "radius_sim": radius_sim, "max_nodes": MX, "m_total": ppos,
        "node_off": np.array(node_off, np.int64), "tet_pad_off": np.array(tet_pad_off, np.int64),
        "N": N, "T": int(tet_pad_off[-1]), "T_real": int(sum(m[2] for m in tile_meta)),
        "x": np.concatenate(pts), "tets": np.concatenate(tets), "ev": np.concatenate(evs),
        "vol": np.concatenate(vols),
        "node_cell": np.repeat(np.arange(nc, dtype=np.uint32), np.diff(np.array(node_off))),
    })
    return h


class BatchSimulation:
    """All cells of ``specs`` on the current device.  ``meshes``: list of (points_um, tets).

    Options (none changes the physics; each cuts bytes per step):
    ``plastic_strain`` keep the accumulated plastic strain (the linear Maxwell law does not read it);
    ``stream_nt`` cache policy of the force pass's streams (None: nontemporal beyond the 256 MB
    memory-side cache); ``sym_state`` viscoplastic state Gd = Fp^-1 Fp^-T - I (6 floats) instead of
    Q = Fp^-1 - I (9);
    ``narrow`` 16-bit tile-node ids and partial positions (when every cell has < 65536 nodes);
    ``shape_from_ref`` the force pass forms Dm^-1 from the reference positions instead of streaming it.

    State on the device: node displacements u from the FP32 reference positions X0 (``ref``), the
    velocity-Verlet state w = a v + f/2, and for the viscoplastic law the plastic deviation Gd
    (``sym_state``, 6 per tetra) or Q (9 per tetra), all zero at the start."""

    def __init__(self, specs, meshes, steps=10000, ncurve=256, tile_k=256, use_graph=True, device_id=0,
                 plastic_strain=False, stream_nt=None, sym_state=True, narrow=True, shape_from_ref=True):
        import torch

        from nofluidx3d_amd import openCL

        openCL.initializeOpenCLObjects(device_id)
        self.specs = list(specs)
        self.steps = steps
        self.ncurve = ncurve
        self.tile_k = K = int(tile_k)
        self.use_graph = use_graph
        self.time = 0
        nc = len(specs)
        self.ncells = nc
        viscous = [s.viscosity_si is not None for s in specs]
        if any(viscous) and not all(viscous):
            raise ValueError("a batch is either all elastic or all viscoplastic")
        self.model = _lib.NFX_VISCOPLASTIC if viscous[0] else _lib.NFX_MR
        h = build_host_arrays(specs, meshes, steps, K)
        self._meshes = meshes
        self.cells_host, self.units, self.radius_sim = h["cells"], h["units"], h["radius_sim"]
        self.node_off, self.tet_pad_off = h["node_off"], h["tet_pad_off"]
        self.N, self.T, self.T_real, MX = h["N"], h["T"], h["T_real"], h["max_nodes"]
        self.max_nodes = MX

        dev = torch.device("cuda", device_id)
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
        u32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to(dev)
        u16 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.uint16).view(np.int16)).to(dev)
        self.ref = f32(h["x"].T.reshape(-1))
        self.x = torch.zeros(3 * self.N, dtype=torch.float32, device=dev)  # displacements
        self.v = torch.zeros(3 * self.N, dtype=torch.float32, device=dev)
        self.fold = torch.zeros(3 * self.N, dtype=torch.float32, device=dev)
        self.tetras = u32(h["tets"].T.reshape(-1))
        ev = torch.from_numpy(np.ascontiguousarray(h["ev"].T.reshape(-1))).to(dev)
        vol = torch.from_numpy(np.ascontiguousarray(h["vol"])).to(dev)
        self.dminv = torch.empty(9 * self.T, dtype=torch.float32, device=dev)
        self.v0 = torch.empty(self.T, dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        hh = openCL.handle()
        _lib.call("nfx_precompute_f64", hh, ev.data_ptr(), vol.data_ptr(), self.T, self.dminv.data_ptr(),
                  self.v0.data_ptr())
        _lib.call("nfx_stream_sync", hh)
        del ev, vol
        self.sym_state = bool(sym_state) and self.model == _lib.NFX_VISCOPLASTIC
        if self.model == _lib.NFX_VISCOPLASTIC:
            # the plastic deviation, zero while Fp = I: by default its symmetric form Gd (the stress,
            # flow and forces depend on Gd alone, and the batch never reports Fp, so a third of the
            # state traffic is saved), else Q
            self.state = torch.zeros((6 if self.sym_state else 9) * self.T, dtype=torch.float32, device=dev)
            # the linear Maxwell law does not read the accumulated plastic strain: kept on request
            self.eps = torch.zeros(self.T, dtype=torch.float32, device=dev) if plastic_strain else None
            self.plastic_strain = self.eps is not None
        else:
            self.state = self.eps = None
            self.plastic_strain = False
        self.t_count, self.t_nodes, self.t_lend = u32(h["tile_count"]), u32(h["tile_nodes"]), u16(h["tile_lend"])
        self.t_corners, self.t_dst = u16(h["tile_corners"]), u32(h["tile_dst"])
        self.t_ntet, self.t_cell = u16(h["tile_ntet"]), u32(h["tile_cell"])
        # per-tile plastic-activity flags: the state starts at zero everywhere
        self.t_active = torch.zeros(len(h["tile_ntet"]), dtype=torch.int32, device=dev) if self.state is not None else None
        # 16-bit tile-node metadata (absent when a cell has 65536 nodes or more)
        self.t_narrow = (u16(h["tile_nodes16"]), u16(h["tile_dst16"]), u32(h["tile_node_base"]),
                         u32(h["tile_dst_base"])) if narrow and "tile_nodes16" in h else None
        self.node_poff, self.node_cell = u32(h["node_poff"]), u32(h["node_cell"])
        self.cell_node_off = u32(self.node_off)
        self.partials = torch.zeros(3 * h["m_total"], dtype=torch.float32, device=dev)
        self.cells = torch.from_numpy(self.cells_host.view(np.uint8).copy()).to(dev)
        self.curves = torch.zeros(nc * ncurve, dtype=torch.float64, device=dev)
        # the initial state, kept on the device for reset(): positions and the viscoplastic state
        # the recentred FP64 positions and each cell's |p[1].x| (sim_points), kept on the device so that
        # reset() rebuilds the geometry there (RadiusSIM, hence the scale, changes with nu)
        p_rel = np.concatenate([np.asarray(P, np.float64) - np.asarray(P, np.float64)[0] for P, _ in meshes])
        self.p_rel = torch.from_numpy(np.ascontiguousarray(p_rel.T.reshape(-1))).to(dev)
        self.p_abs1 = np.array([abs(float(P[1][0]) - float(P[0][0])) for P, _ in meshes])
        torch.cuda.synchronize(dev)

        p = _lib.Plan()
        p.N, p.T = self.N, self.T
        p.tetras, p.dminv, p.v0 = self.tetras.data_ptr(), self.dminv.data_ptr(), self.v0.data_ptr()
        p.ref, p.disp = self.ref.data_ptr(), self.x.data_ptr()
        p.vel, p.fold = self.v.data_ptr(), self.fold.data_ptr()
        p.force = 0
        p.model = self.model
        if self.state is not None:
            p.vp_state = self.state.data_ptr()
            p.epsbar = self.eps.data_ptr() if self.eps is not None else 0
            p.vp.flow_exponent, p.vp.hardening_exponent = 1.0, 0.0  # every cell: linear Maxwell flow
            p.vp_sym = 1 if self.sym_state else 0
            # the flags pay where cells have a yield offset (tiles below it never flow); with none,
            # every stressed tetra flows from the first steps and the flags are overhead
            p.vp_active = self.t_active.data_ptr() if (self.cells_host["vp_offset"] > 0).any() else 0
        p.stream_nt = default_stream_nt(self.T) if stream_nt is None else int(stream_nt)
        self.shape_from_ref = bool(shape_from_ref)
        p.shape_from_ref = 1 if shape_from_ref else 0
        p.scatter = _lib.SCATTER_TILED
        p.tile_k, p.tile_max_nodes = K, MX
        p.tile_count, p.tile_nodes, p.tile_lend = (t.data_ptr() for t in (self.t_count, self.t_nodes, self.t_lend))
        p.tile_corners, p.tile_dst = self.t_corners.data_ptr(), self.t_dst.data_ptr()
        p.node_off, p.partials = self.node_poff.data_ptr(), self.partials.data_ptr()
        p.substrate, p.indenter = 1, _lib.INDENTER_SPHERE
        self.plan = p
        b = Batch()
        b.ncells = nc
        b.tile_ntet, b.tile_cell = self.t_ntet.data_ptr(), self.t_cell.data_ptr()
        b.node_cell, b.cell_node_off, b.cells = self.node_cell.data_ptr(), self.cell_node_off.data_ptr(), \
            self.cells.data_ptr()
        b.curves, b.ncurve, b.curve_steps = self.curves.data_ptr(), ncurve, steps
        if self.t_narrow is not None:
            b.tile_nodes16, b.tile_dst16, b.tile_node_base, b.tile_dst_base = (t.data_ptr() for t in self.t_narrow)
        self.batch = b

    def reset(self, specs):
        """Start over with new physical parameters on the same meshes: only E, nu, viscosity, yield,
        tip radius, speed and indentation depth may change.  Topology, tiles and reference shapes
        are kept; the geometry in simulation units (RadiusSIM depends on nu) is rebuilt on the device
        from the recentred FP64 positions, exactly as construction built it."""
        import torch

        from nofluidx3d_amd import openCL

        if len(specs) != self.ncells:
            raise ValueError(f"reset needs one spec per cell ({self.ncells}), got {len(specs)}")
        for a, b in zip(specs, self.specs):
            if (a.radius_um, a.spacing_um, a.seed) != (b.radius_um, b.spacing_um, b.seed):
                raise ValueError("reset keeps the meshes: radius, spacing and seed must not change")
        if (specs[0].viscosity_si is not None) != (self.model == _lib.NFX_VISCOPLASTIC):
            raise ValueError("reset keeps the material model")
        cells, units, radius_sim = cell_table(specs, self.steps)
        self.specs = list(specs)
        self.cells_host, self.units, self.radius_sim = cells, units, radius_sim
        self.sync()  # the library's last steps have finished with these buffers
        dev = self.x.device
        scale = torch.from_numpy(radius_sim / self.p_abs1).to(dev)
        torch.cuda.synchronize(dev)
        hh = openCL.handle()
        _lib.call("nfx_batch_geometry", hh, self.p_rel.data_ptr(), scale.data_ptr(), self.node_cell.data_ptr(),
                  self.tetras.data_ptr(), self.t_ntet.data_ptr(), self.tile_k, self.N, self.T, self.ref.data_ptr(),
                  self.dminv.data_ptr(), self.v0.data_ptr())
        _lib.call("nfx_stream_sync", hh)
        if self.state is not None:
            self.state.zero_()
            self.t_active.zero_()
        self.x.zero_()
        self.v.zero_()
        self.fold.zero_()
        if self.eps is not None:
            self.eps.zero_()
        self.cells.copy_(torch.from_numpy(self.cells_host.view(np.uint8).copy()))
        self.curves.zero_()
        torch.cuda.synchronize(dev)
        self.time = 0

    def features(self):
        """Surrogate inputs of every cell (``surrogate_inputs``)."""
        return surrogate_inputs([s.E_si for s in self.specs], [s.nu for s in self.specs],
                                [s.viscosity_si or 1.0 for s in self.specs], [s.tip_radius_um for s in self.specs],
                                [s.speed_um_s for s in self.specs], [s.radius_um for s in self.specs],
                                [s.spacing_um for s in self.specs])

    def advance(self, nsteps):
        from nofluidx3d_amd import openCL

        _lib.call("nfx_batch_run_steps", openCL.handle(), self.plan, self.batch, self.time, nsteps,
                  1 if self.use_graph else 0)
        self.time += nsteps

    def sync(self):
        import torch

        from nofluidx3d_amd import openCL

        _lib.call("nfx_stream_sync", openCL.handle())
        torch.cuda.synchronize()

    def time_passes(self, nrep=20):
        from nofluidx3d_amd import openCL

        f, n, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _lib.call("nfx_batch_time_passes", openCL.handle(), self.plan, self.batch, self.time, nrep, ctypes.byref(f),
                  ctypes.byref(n), ctypes.byref(c))
        self.time += nrep
        return f.value, n.value, c.value

    def curves_sim(self):
        """(ncells, ncurve) indenter force in simulation units."""
        self.sync()
        return self.curves.cpu().numpy().reshape(self.ncells, self.ncurve)

    def curves_normalized(self):
        """(ncells, ncurve) F / (E R^2)."""
        return self.curves_sim() / (E_SIM * self.radius_sim[:, None] ** 2)

    def positions(self, i):
        """cell i's node positions X0 + u (N_i, 3), summed in float64"""
        self.sync()
        sl = slice(self.node_off[i], self.node_off[i + 1])
        ref = self.ref.cpu().numpy().reshape(3, self.N).T[sl].astype(np.float64)
        u = self.x.cpu().numpy().reshape(3, self.N).T[sl].astype(np.float64)
        return ref + u

    def single_cell_config(self, i, outdir="/tmp/nfx_batch_cell/"):
        """The reference-format config of cell i (for the single-cell Simulation path)."""
        s, u = self.specs[i], self.units[i]
        cfg = {
            "gpu": 0,
            "RhoSI": u["rho"],
            "SimulationTimeSI": u["T0"] * self.steps,
            "VTKFramesTotal": 1,
            "OutputDir": outdir,
            "InitialDistance": 0.0,
            "SubstrateDistance": 0.0,
            "MoveTimeSI": u["T0"] * self.steps,
            "VelocitySI": s.speed_um_s * 1e-6,
            "SphereRadiusSI": s.tip_radius_um * 1e-6,
            "CELL": {"RadiusSI": s.radius_um * 1e-6, "RadiusSIM": u["RadiusSIM"], "YoungsModulusSI": s.E_si,
                     "PoissonRatio": s.nu, "MooneyRivlinRatio": 1.0, "model": "NEO_HOOKEAN", "InitVTK": ""},
        }
        if s.viscosity_si is not None:
            cfg["CELL"]["ViscositySI"] = s.viscosity_si
            cfg["CELL"]["YieldStressSI"] = s.yield_si
        return cfg



def cell_groups(meshes, max_tetra):
    """Consecutive cell ranges [a, b) whose tetra counts sum to at most ``max_tetra`` (a cell larger
    than that forms a group of its own)."""
    ranges, a, acc = [], 0, 0
    for i, (_, T) in enumerate(meshes):
        if i > a and acc + len(T) > max_tetra:
            ranges.append((a, i))
            a, acc = i, 0
        acc += len(T)
    ranges.append((a, len(meshes)))
    return ranges


class CellGroups:
    """A population of cells advanced group by group: consecutive cells whose tetra counts sum to
    at most ``max_tetra`` form one BatchSimulation, and each group runs its steps before the next
    one starts.  Cells are independent and tiles never span two cells, so every cell's trajectory
    is bitwise the one it has in a single batch.  A group whose per-step working set fits the
    256 MB memory-side cache is served from it from step to step under the default cache policy
    (``autotune.group_tetra``), which beats streaming the whole population past it each step.
    ``max_tetra=None``: one group (a single batch).  Other keyword arguments go to every group's
    BatchSimulation; ``stream_nt`` defaults per group (the default policy for a group that fits).

    Offers the BatchSimulation interface the benchmark, the surrogate's data generation and the
    fitter use: advance, sync, reset, curves_sim, curves_normalized, features, positions,
    single_cell_config, time_passes, and the per-cell specs, units, radius_sim, cells_host."""

    def __init__(self, specs, meshes, max_tetra=None, **kw):
        self.ranges = cell_groups(meshes, max_tetra) if max_tetra else [(0, len(specs))]
        self.groups = [BatchSimulation(specs[a:b], meshes[a:b], **kw) for a, b in self.ranges]
        g0 = self.groups[0]
        self.steps, self.ncurve, self.tile_k, self.sym_state = g0.steps, g0.ncurve, g0.tile_k, g0.sym_state
        self.shape_from_ref, self.plastic_strain = g0.shape_from_ref, g0.plastic_strain
        self.ncells = len(specs)
        self.T_real = sum(g.T_real for g in self.groups)
        self.N = sum(g.N for g in self.groups)
        self.stream_nt = [int(g.plan.stream_nt) for g in self.groups]
        self._collect()

    def _collect(self):
        self.specs = [s for g in self.groups for s in g.specs]
        self.units = [u for g in self.groups for u in g.units]
        self.radius_sim = np.concatenate([g.radius_sim for g in self.groups])
        self.cells_host = np.concatenate([g.cells_host for g in self.groups])

    def _locate(self, i):
        for (a, b), g in zip(self.ranges, self.groups):
            if a <= i < b:
                return g, i - a
        raise IndexError(i)

    @property
    def time(self):
        return self.groups[0].time

    @property
    def eps(self):
        return self.groups[0].eps

    def advance(self, nsteps):
        for g in self.groups:
            g.advance(nsteps)

    def sync(self):
        for g in self.groups:
            g.sync()

    def reset(self, specs):
        if len(specs) != self.ncells:
            raise ValueError(f"reset needs one spec per cell ({self.ncells}), got {len(specs)}")
        for (a, b), g in zip(self.ranges, self.groups):
            g.reset(specs[a:b])
        self._collect()

    def time_passes(self, nrep=20):
        """per-step ms of the force, node and curve passes, summed over the groups"""
        t = np.array([g.time_passes(nrep) for g in self.groups])
        return tuple(float(v) for v in t.sum(axis=0))

    def curves_sim(self):
        return np.concatenate([g.curves_sim() for g in self.groups])

    def curves_normalized(self):
        return np.concatenate([g.curves_normalized() for g in self.groups])

    def features(self):
        return np.concatenate([g.features() for g in self.groups])

    def positions(self, i):
        g, j = self._locate(i)
        return g.positions(j)

    def single_cell_config(self, i, outdir="/tmp/nfx_batch_cell/"):
        g, j = self._locate(i)
        return g.single_cell_config(j, outdir)

CELL_DTYPE = np.dtype([
    ("mu1", np.float32), ("mu2", np.float32), ("kappa", np.float32),
    ("vp_rate", np.float32), ("vp_yield", np.float32), ("vp_offset", np.float32),
    ("wall", np.float32), ("substrate_k", np.float32),
    ("tip_radius", np.float32), ("tip_k", np.float32),
    ("tip_start", np.float64), ("tip_speed", np.float64), ("tip_move_steps", np.float64),
], align=True)


Batch = _lib.Batch
