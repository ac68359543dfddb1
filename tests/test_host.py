"""CPU tests: the C-ABI library's exports, host-side helpers, mesh generator, VTK I/O, TetraCell,
KernelBuilder and the config unit mappings.  No GPU compute is called."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="nfx.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nfx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from nofluidx3d_amd import _lib

    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), f"libnfx.so does not export {s}"
        assert s in _lib.SIGNATURES, f"no ctypes signature for {s}"


def header_struct_fields(text, name):
    """field names of ``typedef struct { ... } name;`` in the header text"""
    end = text.index(f"}} {name};")
    body = text[text.rindex("typedef struct {", 0, end) + len("typedef struct {"):end]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        parts = [w for w in decl.replace("*", " ").split() if w not in ("const", "struct")]
        if not parts:
            continue
        for n in " ".join(parts[1:]).split(","):
            names.append(re.sub(r"\[.*\]", "", n).strip())
    return names


def test_header_structs_match_ctypes():
    """field order of every struct in include/nfx.h equals its ctypes mirror (and the nfx_cell
    record's numpy dtype)"""
    from nofluidx3d_amd import _lib
    from nofluidx3d_amd.batch import CELL_DTYPE

    text = open(os.path.join(ROOT, "include", "nfx.h")).read()
    mirrors = {"nfx_elastic": _lib.Elastic, "nfx_viscoplastic": _lib.Viscoplastic, "nfx_plan": _lib.Plan,
               "nfx_batch": _lib.Batch, "nfx_viscoelastic": _lib.Viscoelastic, "nfx_mlp_gemm_args": _lib.MlpGemm}
    assert set(re.findall(r"^} (\w+);", text, flags=re.M)) == set(mirrors) | {"nfx_cell"}
    for name, cls in mirrors.items():
        assert header_struct_fields(text, name) == [f[0] for f in cls._fields_], name
    assert header_struct_fields(text, "nfx_cell") == list(CELL_DTYPE.names)


def test_node_csr_host_matches_numpy():
    from nofluidx3d_amd import _lib
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, T = generate_sphere_mesh(3.0, 0.75, seed=5)
    N, nT = len(P), len(T)
    tets = np.ascontiguousarray(T.reshape(-1, order="F").astype(np.uint32))
    off = np.zeros(N + 1, np.uint32)
    slots = np.zeros(4 * nT, np.uint32)
    _lib.call("nfx_build_node_csr_host", tets.ctypes.data_as(ctypes.c_void_p), N, nT,
              off.ctypes.data_as(ctypes.c_void_p), slots.ctypes.data_as(ctypes.c_void_p))
    # brute force: for node n, the ascending list of 4*t + a with T[t, a] == n
    flat_slot = (4 * np.arange(nT)[:, None] + np.arange(4)[None, :]).ravel()
    flat_node = T.ravel()
    for n in range(0, N, 7):
        exp = np.sort(flat_slot[flat_node == n])
        assert np.array_equal(slots[off[n]:off[n + 1]], exp)
    assert off[-1] == 4 * nT
    # a corner index out of range is rejected, not written
    bad = tets.copy()
    bad[5] = N + 3
    rc = _lib.lib().nfx_build_node_csr_host(bad.ctypes.data_as(ctypes.c_void_p), N, nT,
                                            off.ctypes.data_as(ctypes.c_void_p),
                                            slots.ctypes.data_as(ctypes.c_void_p))
    assert rc != 0 and b"references node" in _lib.lib().nfx_last_error()


def test_null_arguments_are_errors_not_crashes():
    from nofluidx3d_amd import _lib

    L = _lib.lib()
    assert L.nfx_stream_sync(None) == -1
    assert L.nfx_tip_force(None, None, None, 10, 1, 0.1, 1.0, 0.0, None) == -1
    assert L.nfx_run_steps(None, None, 0, 1, None, 0) == -1


def test_mesh_generator_invariants():
    from nofluidx3d_amd.mesh import DEFAULT_MIN_QUALITY, generate_sphere_mesh, signed_volumes, tetra_quality

    P, T = generate_sphere_mesh(8.0, 0.8, seed=1)
    assert (len(P), len(T)) == (4169, 25342)  # the seeded small cell
    assert np.array_equal(P[0], [0, 0, 0]) and np.array_equal(P[1], [8.0, 0, 0])
    assert np.all(signed_volumes(P, T) > 0)
    assert tetra_quality(P, T).min() >= DEFAULT_MIN_QUALITY
    assert len(np.unique(T)) == len(P)  # every node used
    # boundary faces (in exactly one tetra) form closed surfaces: every edge lies in an even number
    # of them (the boundary of a boundary is empty mod 2)
    faces = np.sort(np.concatenate([T[:, [1, 2, 3]], T[:, [0, 2, 3]], T[:, [0, 1, 3]], T[:, [0, 1, 2]]]), 1)
    uf, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2
    bnd = uf[cnt == 1]
    edges = np.sort(np.concatenate([bnd[:, [0, 1]], bnd[:, [1, 2]], bnd[:, [0, 2]]]), 1)
    _, ecnt = np.unique(edges, axis=0, return_counts=True)
    assert np.all(ecnt % 2 == 0)
    # deterministic from the seed
    P2, T2 = generate_sphere_mesh(8.0, 0.8, seed=1)
    assert np.array_equal(P, P2) and np.array_equal(T, T2)


@pytest.mark.parametrize("binary", [True, False])
def test_vtk_roundtrip(tmp_path, binary):
    from nofluidx3d_amd.mesh import generate_sphere_mesh
    from nofluidx3d_amd.util import vtk

    P, T = generate_sphere_mesh(2.0, 0.5, seed=3)
    g = vtk.from_tetra(P, T)
    g.cell_data.set_array(np.arange(len(T), dtype=np.float64) * 0.5, "vonMises")
    path = tmp_path / "m.vtk"
    g.save(str(path), binary=binary)
    h = vtk.read(str(path))
    assert np.array_equal(h.points, P)
    assert np.array_equal(h.cells_dict[10], T)
    assert np.array_equal(h.cell_data["vonMises"], g.cell_data["vonMises"])


def test_vtk_reads_version_51_ascii(tmp_path):
    from nofluidx3d_amd.util import vtk

    text = """# vtk DataFile Version 5.1
cell
ASCII
DATASET UNSTRUCTURED_GRID
POINTS 5 float
0 0 0 1 0 0 0 1 0
0 0 1 1 1 1
CELLS 3 8
OFFSETS vtktypeint64
0 4 8
CONNECTIVITY vtktypeint64
0 1 2 3 1 2 3 4
CELL_TYPES 2
10
10
CELL_DATA 2
SCALARS p double 1
LOOKUP_TABLE default
1.5 2.5
"""
    path = tmp_path / "m51.vtk"
    path.write_text(text)
    g = vtk.read(str(path))
    assert g.n_points == 5 and g.n_cells == 2
    assert np.array_equal(g.cells_dict[10], [[0, 1, 2, 3], [1, 2, 3, 4]])
    assert np.allclose(g.cell_data["p"], [1.5, 2.5])


def test_tetracell_recenter_scale_and_reference_quantities(tmp_path):
    import oracle as O
    from nofluidx3d_amd.mesh import generate_sphere_mesh
    from nofluidx3d_amd.TetraCell import TetraCell
    from nofluidx3d_amd.util import vtk

    P, T = generate_sphere_mesh(3.0, 0.6, seed=2)
    shifted = P * 1e-6 + np.array([1e-6, -2e-6, 3e-6])
    path = tmp_path / "cell.vtk"
    vtk.from_tetra(shifted, T).save(str(path))
    c = TetraCell(str(path), radius=15.0)
    assert np.allclose(c.points[0], 0)
    assert abs(c.points[1][0]) == pytest.approx(15.0)
    assert np.allclose(c.points, P * 5.0, atol=1e-9)
    ev, vol = O.edge_vectors_volumes(c.points, c.tetras)
    assert np.allclose(c.edgeVectors, ev) and np.allclose(c.volumes, vol)
    c.addTetraData("vonMises", np.ones(len(T)))
    assert c.mesh.cell_data["vonMises"].shape == (len(T),)
    c.mesh.points = c.points * 1.1
    c.calcVolumeChange()
    assert np.allclose(c.mesh.cell_data["volumeChange"], 1.1**3 - 1)


def test_config_mesh_written_to_init_vtk_reads_back_through_tetracell(tmp_path):
    """configs.write_mesh puts a config's seeded mesh (micrometres) at CELL.InitVTK in metres;
    TetraCell reads it back recentred on node 0 and scaled to RadiusSIM."""
    from nofluidx3d_amd import configs
    from nofluidx3d_amd.TetraCell import TetraCell

    cfg = configs.small_cell(outdir=str(tmp_path) + "/", steps=10, frames=1)
    cfg["MESH"] = {"radius_um": 3.0, "spacing_um": 0.75, "seed": 4}
    cfg["CELL"]["InitVTK"] = str(tmp_path / "mesh" / "cell.vtk")
    P, T = configs.write_mesh(cfg)
    c = TetraCell(cfg["CELL"]["InitVTK"], radius=cfg["CELL"]["RadiusSIM"])
    assert np.array_equal(c.tetras, T)
    scale = cfg["CELL"]["RadiusSIM"] / abs(P[1, 0] - P[0, 0])
    # held at the FP32 values the device keeps: within FP32 rounding of the FP64 recentred mesh
    assert np.allclose(c.points, (P - P[0]) * scale, rtol=0, atol=1e-7 * cfg["CELL"]["RadiusSIM"])
    assert np.array_equal(c.points, c.points.astype(np.float32).astype(np.float64))


def test_cell_groups_partition_consecutive_cells_under_the_budget():
    """batch.cell_groups: consecutive ranges covering every cell once, each within the tetra budget
    unless it is a single cell larger than the budget; a budget above the total gives one group"""
    from nofluidx3d_amd.autotune import MEMORY_SIDE_CACHE_BYTES, RESIDENT_BYTES_PER_TETRA, group_tetra
    from nofluidx3d_amd.batch import cell_groups

    rng = np.random.default_rng(8)
    sizes = rng.integers(100, 5000, size=200)
    sizes[17] = 20000  # a cell above the budget
    meshes = [(None, np.zeros((int(t), 4))) for t in sizes]
    budget = 9000
    r = cell_groups(meshes, budget)
    assert r[0][0] == 0 and r[-1][1] == len(sizes)
    assert all(a < b for a, b in r) and all(r[i][1] == r[i + 1][0] for i in range(len(r) - 1))
    for a, b in r:
        tot = int(sizes[a:b].sum())
        assert tot <= budget or b - a == 1
        if b < len(sizes):
            assert tot + int(sizes[b]) > budget  # groups are filled greedily
    assert cell_groups(meshes, int(sizes.sum())) == [(0, len(sizes))]
    assert group_tetra() == MEMORY_SIDE_CACHE_BYTES // RESIDENT_BYTES_PER_TETRA


def test_split_k_choice_divides_the_batch_and_fills_the_grid():
    """surrogate.mlp.split_k: a power of two <= 64 whose slices are whole multiples of 32 rows of K,
    doubled only while the grid of 128 x 32 tiles stays under 512 workgroups"""
    from nofluidx3d_amd.surrogate.mlp import split_k

    for M, N, K in [(512, 32, 4096), (512, 512, 4096), (256, 512, 4096), (128, 32, 256), (512, 512, 128),
                    (4096, 512, 512), (128, 32, 96)]:
        s = split_k(M, N, K)
        groups = (M // 128) * (N // 32)
        assert s & (s - 1) == 0 and 1 <= s <= 64 and K % (32 * s) == 0, (M, N, K, s)
        if s > 1:
            assert groups * (s // 2) < 512  # doubled only while the grid was short
        assert s == 64 or groups * s >= 512 or K % (64 * s) != 0, (M, N, K, s)
    assert split_k(512, 512, 4096) == 8 and split_k(4096, 512, 512) == 1


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_kernel_builder_defines_render_like_the_reference():
    from nofluidx3d_amd.KernelBuilder import KernelBuilder

    KernelBuilder.define("FLAG_X", a=3, b=True, c=0.5)
    s = KernelBuilder.getDefines()
    assert "#define FLAG_X\n" in s and "#define a 3\n" in s and "#define b true\n" in s and "#define c 0.5\n" in s
    k = KernelBuilder.build("kernels/Interactions/tetra.cl", "calcTetraForce")
    assert k.values["a"] == 3 and "FLAG_X" in k.defines
    with pytest.raises(Exception):
        KernelBuilder.build("kernels/CalcCOM.cl", "CalcCOM")


def test_config_units():
    import oracle as O
    from nofluidx3d_amd import configs

    c = configs.small_cell()
    u = O.units(c["CELL"]["RadiusSI"], c["CELL"]["RadiusSIM"], c["CELL"]["YoungsModulusSI"], c["RhoSI"])
    assert round(c["SimulationTimeSI"] / u["T0"]) == 20000
    assert round(round(c["SimulationTimeSI"] / u["T0"]) / c["VTKFramesTotal"]) == 100
    # tip travels 2 um in the move time
    assert c["VelocitySI"] * c["MoveTimeSI"] == pytest.approx(2e-6)
    L = configs.large_cell()
    cc = L["CELL"]
    mu = cc["YoungsModulusSI"] / (2 * (1 + cc["PoissonRatio"]))
    K = cc["YoungsModulusSI"] / (3 * (1 - 2 * cc["PoissonRatio"]))
    assert mu * cc["MooneyRivlinRatio"] == pytest.approx(300.0)
    assert mu * (1 - cc["MooneyRivlinRatio"]) == pytest.approx(100.0)
    assert K == pytest.approx(5000.0)


@pytest.mark.parametrize("K", [256, 512])
def test_tile_builder_invariants(K):
    from nofluidx3d_amd import _lib
    from nofluidx3d_amd.mesh import generate_sphere_mesh

    P, T = generate_sphere_mesh(3.0, 0.5, seed=6)
    N, nT = len(P), len(T)
    tets = np.ascontiguousarray(T.reshape(-1, order="F").astype(np.uint32))
    vp = ctypes.c_void_p
    ntiles = (nT + K - 1) // K
    count = np.zeros(ntiles, np.uint32)
    m, mx = ctypes.c_int64(), ctypes.c_int32()
    _lib.call("nfx_build_tiles_host", tets.ctypes.data_as(vp), N, nT, K, ctypes.byref(mx), ctypes.byref(m),
              count.ctypes.data_as(vp), None, None, None, None, None)
    MX, M = mx.value, m.value
    assert MX == count.max() and M == count.sum() and MX <= K
    tn, tl = np.zeros(ntiles * MX, np.uint32), np.zeros(ntiles * MX, np.uint16)
    ts, td = np.zeros(4 * K * ntiles, np.uint16), np.zeros(ntiles * MX, np.uint32)
    no = np.zeros(N + 1, np.uint32)
    _lib.call("nfx_build_tiles_host", tets.ctypes.data_as(vp), N, nT, K, ctypes.byref(mx), ctypes.byref(m),
              count.ctypes.data_as(vp), tn.ctypes.data_as(vp), tl.ctypes.data_as(vp), ts.ctypes.data_as(vp),
              td.ctypes.data_as(vp), no.ctypes.data_as(vp))
    owner = -np.ones(M, np.int64)  # which (tile, tile-node) fills each node-ordered partial position
    for c in range(ntiles):
        t0, nt = c * K, min(K, nT - c * K)
        sub = T[t0:t0 + nt]
        nodes = tn[c * MX:c * MX + count[c]]
        assert np.array_equal(nodes, np.unique(sub))  # ascending unique nodes of the tile
        ends = tl[c * MX:c * MX + count[c]].astype(np.int64)
        starts = np.concatenate([[0], ends[:-1]])
        assert ends[-1] == 4 * nt
        rec = ts[4 * K * c:4 * K * c + 4 * nt].astype(np.int64)
        local, rank = rec & 0xFF, rec >> 8
        # every corner names its node, and its places (run start + rank) enumerate the sorted list
        assert np.all(nodes[local] == sub.reshape(-1))
        place = starts[local] + rank
        assert sorted(place) == list(range(4 * nt))
        for i, (b, e) in enumerate(zip(starts, ends)):
            slots = np.nonzero(local == i)[0]  # this node's corner slots, ascending
            assert len(slots) == e - b
            assert np.array_equal(place[slots], np.arange(b, e))  # slot order within the run
            d = td[c * MX + i]
            assert no[nodes[i]] <= d < no[nodes[i] + 1]
            assert owner[d] == -1
            owner[d] = c
        assert np.all(ts[4 * K * c + 4 * nt:4 * K * (c + 1)] == 0)  # padding slots
        assert np.all(tn[c * MX + count[c]:(c + 1) * MX] < N)  # padding stays a valid node id
    assert np.all(owner >= 0) and no[-1] == M
    # within a node, partial positions follow tile order
    for n in range(N):
        assert np.all(np.diff(owner[no[n]:no[n + 1]]) > 0)
    # the node-fused step's per-tile-node partial ranges, and its storing tile: for every node exactly
    # one tile-node, the node's first tile, has its partial at the start of the range
    from nofluidx3d_amd.Simulation import tile_ranges

    rg = tile_ranges(tn, count, no, MX).reshape(ntiles, MX, 2)
    stores = np.zeros(N, np.int64)
    for c in range(ntiles):
        for i in range(MX):
            if i >= count[c]:
                assert rg[c, i, 0] == 0 and rg[c, i, 1] == 0
                continue
            n = tn[c * MX + i]
            assert rg[c, i, 0] == no[n] and rg[c, i, 1] == no[n + 1]
            stores[n] += td[c * MX + i] == rg[c, i, 0]
    assert np.all(stores == 1) and np.all(np.diff(no) > 0)  # every node has a partial (node-fused form)


def test_step_form_default_by_mesh_size():
    from nofluidx3d_amd import autotune

    assert autotune.default_node_fused(25_342) and autotune.default_node_fused(autotune.NODE_FUSED_MAX_TETRA)
    assert not autotune.default_node_fused(autotune.NODE_FUSED_MAX_TETRA + 1)
    assert not autotune.default_node_fused(3_377_961)


def test_cell_table_layout_matches_header():
    from nofluidx3d_amd import _lib
    from nofluidx3d_amd.batch import CELL_DTYPE

    text = open(os.path.join(ROOT, "include", "nfx.h")).read()
    body = text[text.index("typedef struct {\n    float mu1, mu2, kappa;"):text.index("} nfx_cell;")]
    names = re.findall(r"\b([a-z_0-9]+)\s*[,;]", body.replace("typedef struct {", ""))
    assert list(CELL_DTYPE.names) == names
    assert CELL_DTYPE.itemsize == 64 and CELL_DTYPE.fields["tip_start"][1] == 40
    body = text[text.index("typedef struct {\n    int32_t ncells;"):text.index("} nfx_batch;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = [d.replace("*", " ").split()[-1] for d in body.replace("typedef struct {", "").split(";") if d.strip()]
    assert names == [f[0] for f in _lib.Batch._fields_]


def test_batch_sampling_respects_ranges():
    from nofluidx3d_amd.batch import estimated_tetra, sample_batch

    specs, stats = sample_batch(64, seed0=1000)
    assert len(specs) == 64 and len({s.seed for s in specs}) == 64
    for s in specs:
        assert 4 <= s.radius_um <= 10 and 0.3 <= s.spacing_um <= 0.6
        assert 2000 <= estimated_tetra(s.radius_um, s.spacing_um) <= 60000
        assert 100 <= s.E_si <= 1e4 and 0.3 <= s.nu <= 0.49 and 1 <= s.viscosity_si <= 1000
        assert 1 <= s.tip_radius_um <= 5 and 0.5 <= s.speed_um_s <= 50
        assert s.relaxation_per_step(10000) <= 0.1
    again, _ = sample_batch(64, seed0=1000)
    assert again == specs  # seeded


def test_batch_host_arrays_indexing_is_in_bounds():
    """Every index the batched kernels read from the host-built arrays stays inside its array."""
    from nofluidx3d_amd.batch import build_host_arrays, make_meshes, sample_batch

    specs, _ = sample_batch(6, seed0=77, t_range=(2000, 9000))
    meshes = make_meshes(specs, workers=1, cache=None)
    for K in (256, 512):
        h = build_host_arrays(specs, meshes, 1000, K)
        N, T, MX, M = h["N"], h["T"], h["max_nodes"], h["m_total"]
        ntiles = len(h["tile_count"])
        assert T == ntiles * K and MX <= K
        assert h["tile_nodes"].max() < N and h["tile_dst"].max() < M
        assert np.all(h["tile_ntet"] >= 1) and np.all(h["tile_ntet"] <= K)
        assert np.all(h["tile_count"] <= MX) and h["tile_count"].sum() == M
        assert h["node_poff"][-1] == M and np.all(np.diff(h["node_poff"].astype(np.int64)) >= 1)
        assert h["tets"].max() < N and len(h["tets"]) == T
        # every tile's node ids lie inside its own cell's node range
        for c in range(ntiles):
            cell = h["tile_cell"][c]
            nodes = h["tile_nodes"][c * MX:c * MX + h["tile_count"][c]]
            assert np.all((nodes >= h["node_off"][cell]) & (nodes < h["node_off"][cell + 1]))
            lend = h["tile_lend"][c * MX:(c + 1) * MX]
            assert lend.max() <= 4 * h["tile_ntet"][c]
        # tile-node partial positions are a permutation of 0..M-1
        used = np.concatenate([h["tile_dst"][c * MX:c * MX + h["tile_count"][c]] for c in range(ntiles)])
        assert np.array_equal(np.sort(used), np.arange(M))
        assert len(h["node_cell"]) == N and h["node_cell"].max() == len(specs) - 1


def test_symmetric_viscoplastic_state_update_identity():
    """The batch's viscoplastic state G = S S^T (S = Dm^-1 Fp^-1) evolves on its own: the update
    G' = G + r^-1 (-N Be - Be N + N Be N) r^-T with Be = r G r^T and N = (I + Dp)^-1 Dp equals
    S' S'^T of the shape-state update S' = S - r^-1 (I + Dp)^-1 Dp (r S), and Be = Fe Fe^T, for
    random edges, states and trace-free symmetric Dp (FP64)."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        r = rng.normal(size=(3, 3)) * 0.2 + np.eye(3)  # well-conditioned edge matrices
        S = np.linalg.inv(r) @ (np.eye(3) + 0.2 * rng.normal(size=(3, 3)))
        A = rng.normal(size=(3, 3))
        D = (A + A.T) / 2
        D -= np.trace(D) / 3 * np.eye(3)
        D *= rng.uniform(1e-4, 0.2) / np.abs(D).max()
        Fe = r @ S
        Sp = S - np.linalg.inv(r) @ np.linalg.inv(np.eye(3) + D) @ D @ Fe
        G = S @ S.T
        Be = r @ G @ r.T
        N = np.linalg.inv(np.eye(3) + D) @ D
        ri = np.linalg.inv(r)
        Gp = G + ri @ (-N @ Be - Be @ N + N @ Be @ N) @ ri.T
        assert np.allclose(Be, Fe @ Fe.T, rtol=0, atol=1e-11 * np.abs(Be).max())
        assert np.allclose(Gp, Sp @ Sp.T, rtol=0, atol=1e-11 * np.abs(Gp).max())
        # a dropped term fails: without N Be N the update is off by O(|Dp|^2)
        Gx = G + ri @ (-N @ Be - Be @ N) @ ri.T
        assert not np.allclose(Gx, Sp @ Sp.T, rtol=0, atol=1e-11 * np.abs(Gp).max())
    # no flow: the increment vanishes exactly
    assert np.array_equal(np.linalg.inv(np.eye(3) + np.zeros((3, 3))) @ np.zeros((3, 3)), np.zeros((3, 3)))


def test_narrow_tile_arrays_round_trip():
    """The batch's 16-bit tile-node metadata (offsets from each tile's first entry) gives back the
    32-bit ids and partial positions of every real tile-node, and is refused when an offset would
    not fit 16 bits."""
    from nofluidx3d_amd.batch import narrow_tile_arrays

    rng = np.random.default_rng(4)
    ntiles, MX = 40, 24
    count = rng.integers(5, MX + 1, ntiles).astype(np.uint32)
    nodes = np.zeros((ntiles, MX), np.uint32)
    dst = np.zeros((ntiles, MX), np.uint32)
    for c in range(ntiles):
        base = rng.integers(0, 10 ** 6)
        nodes[c, :count[c]] = base + np.sort(rng.choice(60000, count[c], replace=False))
        nodes[c, count[c]:] = nodes[c, 0]  # padding: a valid id
        dst[c, :count[c]] = 3 * base + np.sort(rng.choice(60000, count[c], replace=False))
    h = narrow_tile_arrays(nodes.reshape(-1), dst.reshape(-1), count, MX)
    n16, d16 = h["tile_nodes16"].reshape(ntiles, MX), h["tile_dst16"].reshape(ntiles, MX)
    for c in range(ntiles):
        k = count[c]
        assert np.array_equal(h["tile_node_base"][c] + n16[c, :k].astype(np.uint32), nodes[c, :k])
        assert np.array_equal(h["tile_dst_base"][c] + d16[c, :k].astype(np.uint32), dst[c, :k])
    nodes[3, count[3] - 1] = nodes[3, 0] + 70000  # an offset beyond 16 bits
    assert narrow_tile_arrays(nodes.reshape(-1), dst.reshape(-1), count, MX) == {}
