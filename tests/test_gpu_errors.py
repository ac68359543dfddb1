"""Argument errors at the C ABI on a live context: a bad size, option or pointer is refused with a
negative code and a message naming it, before anything is launched, and the state is left as it
was (a valid step after the refusals equals the same step of an untouched twin, bitwise)."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd import openCL

    openCL.initializeOpenCLObjects(0)
    return openCL.handle()


def small_sim():
    from nofluidx3d_amd import configs
    from nofluidx3d_amd.Simulation import Simulation
    from nofluidx3d_amd.util import vtk

    cfg = configs.small_cell(outdir="/tmp/nfx_test_errors/", steps=40, frames=2)
    cfg["MESH"] = {"radius_um": 4.0, "spacing_um": 0.8, "seed": 1}
    cfg["CELL"]["RadiusSI"] = 4e-6
    cfg["CELL"]["RadiusSIM"] = 10.0
    cfg["SphereRadiusSI"] = 1.6e-6
    cfg["InitialDistance"] = -0.5
    P, T = configs.make_mesh(cfg)
    sim = Simulation(cfg, scatter="tiled", mesh=vtk.from_tetra(P, T), verbose=False, tile_k=256)
    sim.setInteractionSubstrate()
    sim.setInteractionSphere()
    sim.setInteractionTetra()
    sim.setInteractionVelocityVerlet()
    return sim


def positions(sim):
    import torch

    from nofluidx3d_amd import _lib, openCL

    _lib.call("nfx_stream_sync", openCL.handle())
    torch.cuda.synchronize()
    return sim.displacement.cpu().numpy().copy()


def refused(name, *args, expect):
    from nofluidx3d_amd import _lib

    assert len(args) == len(_lib.SIGNATURES[name][1]), f"{name}: argument count"
    rc = getattr(_lib.lib(), name)(*args)
    msg = _lib.lib().nfx_last_error().decode()
    assert rc < 0, f"{name} accepted a bad argument"
    assert expect in msg, msg
    return rc


def copy_plan(p, **changes):
    from nofluidx3d_amd import _lib

    q = _lib.Plan()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(_lib.Plan))
    for k, v in changes.items():
        setattr(q, k, v)
    return q


PLAN_CASES = [
    ({"N": 0}, "node count 0 out of range"),
    ({"T": 1 << 30}, "tetra count"),
    ({"model": 99}, "unknown model 99"),
    ({"disp": None}, "NULL mesh/state pointer"),
    ({"ref": None}, "NULL mesh/state pointer"),
    ({"vp_sym": 1}, "symmetric state (vp_sym) is for the tiled viscoplastic force only"),
    ({"scatter": 7}, "unknown scatter 7"),
    ({"tile_k": 100}, "tile_k must be 128, 256 or 512"),
    ({"tile_max_nodes": 300}, "tile_max_nodes 300 exceeds"),
    ({"tile_corners": None}, "tiled scatter needs the tile arrays"),
    ({"indenter": 3}, "unknown indenter 3"),
]


def test_bad_plans_are_refused_and_leave_the_state(ctx):
    from nofluidx3d_amd import _lib

    a, b = small_sim(), small_sim()
    a.advance(5)
    b.advance(5)
    p = a.plan()
    assert p is not None and p.scatter == _lib.SCATTER_TILED
    before = positions(a)
    for changes, expect in PLAN_CASES:
        q = copy_plan(p, **changes)
        for use_graph in (0, 1):
            refused("nfx_run_steps", ctx, ctypes.byref(q), 5, 3, None, use_graph, expect=expect)
    refused("nfx_run_steps", ctx, ctypes.byref(p), 5, -1, None, 0, expect="nsteps < 0")
    # the node-fused arrays (this small cell takes the node-fused form): incomplete, aliasing the
    # plan's own buffers, or with a law the fused form does not cover
    assert p.disp_b and p.tile_range
    for changes, expect in (({"tile_range": None}, "are set together"), ({"vel_b": None}, "are set together"),
                            ({"disp_b": p.disp}, "must differ from disp, vel and partials"),
                            ({"model": _lib.NFX_LINEAR_ELASTIC}, "node-fused steps need the tiled scatter")):
        q = copy_plan(p, **changes)
        refused("nfx_run_steps", ctx, ctypes.byref(q), 5, 3, None, 1, expect=expect)
    assert np.array_equal(positions(a), before)
    a.advance(7)
    b.advance(7)
    assert np.array_equal(positions(a), positions(b))


def test_bad_sizes_and_options_of_other_entry_points(ctx):
    import torch

    from nofluidx3d_amd import _lib

    buf = torch.zeros(4096, dtype=torch.float64, device="cuda:0")
    ptr = buf.data_ptr()
    refused("nfx_precompute_f64", ctx, ptr, ptr, 0, ptr, ptr, expect="tetra count 0 out of range")
    refused("nfx_batch_geometry", ctx, ptr, ptr, ptr, ptr, ptr, 100, 100, 500, ptr, ptr, ptr,
            expect="tile_k must be 128, 256 or 512")
    refused("nfx_batch_geometry", ctx, ptr, ptr, ptr, ptr, ptr, 256, 100, 1000, ptr, ptr, ptr,
            expect="multiple of tile_k")
    refused("nfx_batch_geometry", ctx, ptr, ptr, ptr, ptr, ptr, 256, 0, 512, ptr, ptr, ptr,
            expect="node count 0 out of range")
    g = _lib.MlpGemm()
    g.epilogue, g.M, g.N, g.K = _lib.NFX_MLP_STORE, 100, 64, 64  # M not a multiple of 128
    g.A, g.B, g.C, g.lda, g.ldb = ptr, ptr, ptr, 64, 64
    refused("nfx_mlp_gemm", ctx, ctypes.byref(g), expect="M must be a multiple of 128")
    g.M, g.epilogue = 128, 9
    refused("nfx_mlp_gemm", ctx, ctypes.byref(g), expect="unknown epilogue 9")
    out = ctypes.c_void_p()
    n = ctypes.c_int()
    _lib.call("nfx_device_count", ctypes.byref(n))
    refused("nfx_init", n.value + 3, None, ctypes.byref(out), expect=f"device {n.value + 3} not found")
    assert not out.value
    _lib.call("nfx_stream_sync", ctx)  # the context still works
