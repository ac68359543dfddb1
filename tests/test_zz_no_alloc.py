"""The allocation-free tip-force reduction (nfx_tip_force_noalloc, scratch in a fixed per-GPU array of
the library) against the context entry point nfx_tip_force on the same inputs: bitwise equal, on
two calls in a row (the second with a smaller input, which must not see the first call's scratch),
for displacement and absolute positions and both indenters; and its argument checks."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd import openCL

    openCL.initializeOpenCLObjects(0)
    return openCL.handle()


def _sync(ctx):
    import torch

    from nofluidx3d_amd import _lib

    _lib.call("nfx_stream_sync", ctx)
    torch.cuda.synchronize()


def test_noalloc_tip_force_equals_the_context_entry_point(ctx):
    import torch

    from nofluidx3d_amd import _lib

    rng = np.random.default_rng(7)
    out_ctx = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    out_free = torch.full((2,), -1.0, dtype=torch.float64, device="cuda:0")
    for N in (300000, 1000):  # the second call smaller than the first
        X0 = rng.normal(scale=6.0, size=3 * N).astype(np.float32)
        u = rng.normal(scale=0.5, size=3 * N).astype(np.float32)
        X0d, ud = torch.from_numpy(X0).cuda(), torch.from_numpy(u).cuda()
        xd = torch.from_numpy(X0 + u).cuda()
        x64 = X0.astype(np.float64) + u.astype(np.float64)
        for pts, ref, xs in ((ud, X0d, x64), (xd, None, (X0 + u).astype(np.float64))):
            for ind, R, pos in ((1, 8.0, 9.0), (2, 0.0, 6.0)):
                refp = None if ref is None else ref.data_ptr()
                _lib.call("nfx_tip_force", ctx, pts.data_ptr(), refp, N, ind, 0.1, R, pos, out_ctx.data_ptr())
                _sync(ctx)
                _lib.call("nfx_tip_force_noalloc", pts.data_ptr(), refp, N, ind, 0.1, R, pos, out_free.data_ptr(),
                          None)
                _sync(ctx)
                a, b = out_ctx.cpu().numpy(), out_free.cpu().numpy()
                assert np.array_equal(a, b), (N, ind, ref is None, a, b)
                f, m = O.tip_force(np.ascontiguousarray(xs), ind, 0.1, R, pos)
                assert a[0] == pytest.approx(f, rel=1e-12) and a[1] == pytest.approx(m, rel=1e-12)
                assert f != 0.0


def test_noalloc_tip_force_refuses_bad_arguments(ctx):
    import torch

    from nofluidx3d_amd import _lib

    buf = torch.zeros(64, dtype=torch.float64, device="cuda:0")
    L = _lib.lib()
    for N, ind, expect in ((0, 1, b"node count 0 out of range"), (1 << 31, 1, b"out of range"),
                           (10, 3, b"unknown indenter 3")):
        rc = L.nfx_tip_force_noalloc(buf.data_ptr(), None, N, ind, 0.1, 1.0, 0.0, buf.data_ptr(), None)
        assert rc < 0 and expect in L.nfx_last_error()
    assert L.nfx_tip_force_noalloc(None, None, 10, 1, 0.1, 1.0, 0.0, buf.data_ptr(), None) < 0
