"""GPU tests of the MLP surrogate kernels (BF16 operands on MFMA, FP32 accumulation).

References: the GEMM against an FP64 product of the same BF16-rounded operands (tolerance 1e-5
relative: FP32 accumulation); forward/gradients against the FP64 numpy oracle run on the
BF16-rounded weights and inputs (tolerance 3e-2 relative in norm: BF16 activations and gradients);
the fused Adam update against the oracle's Adam on the kernels' own gradients (1e-6).
"""

import numpy as np
import pytest

import oracle.mlp as OM

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(11)


@pytest.fixture(scope="module")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nofluidx3d_amd import openCL

    openCL.initializeOpenCLObjects(0)
    return openCL.handle()


def bf(a):
    import torch

    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16)


def bf_round(a):
    return bf(a).float().numpy().astype(np.float64)


@pytest.mark.parametrize("M,N,K", [(128, 32, 32), (256, 96, 64), (512, 512, 512), (256, 64, 4096),
                                   (16384, 256, 512)])  # the last takes the LDS-tiled kernel (256 tiles of 128 x 128)
def test_gemm_nt_matches_fp64_product(ctx, M, N, K):
    import torch

    from nofluidx3d_amd import _lib
    from nofluidx3d_amd.surrogate.mlp import MLP

    A, B = RNG.normal(size=(M, K)), RNG.normal(size=(N, K))
    bias = RNG.normal(size=N)
    Ad, Bd = bf(A).cuda(), bf(B).cuda()
    C = torch.empty(M, N, dtype=torch.float32, device="cuda:0")
    bd = torch.tensor(bias, dtype=torch.float32, device="cuda:0")
    MLP._gemm(_lib.NFX_MLP_STORE, M, N, K, Ad, K, Bd, K, bias=bd, C=C)
    ref = bf_round(A) @ bf_round(B).T + bias.astype(np.float32)
    out = C.cpu().numpy()
    assert np.max(np.abs(out - ref)) < 1e-5 * np.max(np.abs(ref)) + 1e-4
    # FWD_GELU epilogue: Z, BF16 H and its exact transpose
    Z = torch.empty(M, N, dtype=torch.float32, device="cuda:0")
    H = torch.empty(M, N, dtype=torch.bfloat16, device="cuda:0")
    HT = torch.empty(N, M, dtype=torch.bfloat16, device="cuda:0")
    MLP._gemm(_lib.NFX_MLP_FWD_GELU, M, N, K, Ad, K, Bd, K, bias=bd, C=Z, H=H, HT=HT)
    z = Z.cpu().numpy()
    assert np.allclose(z, out, rtol=0, atol=1e-6 * np.max(np.abs(out)))
    h = H.float().cpu().numpy()
    assert np.max(np.abs(h - OM.gelu(z))) <= 2 ** -7 * np.max(np.abs(OM.gelu(z))) + 1e-6
    assert torch.equal(H.t().contiguous(), HT)
    # BWD_GELU epilogue: dZ = acc * gelu'(Z)
    dZ = torch.empty(M, N, dtype=torch.bfloat16, device="cuda:0")
    dZT = torch.empty(N, M, dtype=torch.bfloat16, device="cuda:0")
    MLP._gemm(_lib.NFX_MLP_BWD_GELU, M, N, K, Ad, K, Bd, K, Z=Z, H=dZ, HT=dZT)
    exp = (bf_round(A) @ bf_round(B).T) * OM.gelu_grad(z)
    got = dZ.float().cpu().numpy()
    assert np.max(np.abs(got - exp)) <= 2 ** -7 * np.max(np.abs(exp)) + 1e-5
    assert torch.equal(dZ.t().contiguous(), dZT)


@pytest.mark.parametrize("M,N,K,S", [(128, 32, 4096, 64), (512, 512, 4096, 8), (256, 64, 1024, 4)])
def test_split_k_matches_one_pass(ctx, M, N, K, S):
    """split-K (K slices accumulated into a workspace, summed in slice order, then the epilogue)
    against the one-pass GEMM: STORE within FP32 rounding of the FP64 product; the fused Adam
    update of the same master weights within 1e-6, its BF16 copies exact transposes"""
    import math

    import torch

    from nofluidx3d_amd import _lib
    from nofluidx3d_amd.surrogate.mlp import MLP

    A, B = RNG.normal(size=(M, K)), RNG.normal(size=(N, K))
    Ad, Bd = bf(A).cuda(), bf(B).cuda()
    ws = torch.empty(S * M * N, dtype=torch.float32, device="cuda:0")
    ref = bf_round(A) @ bf_round(B).T
    outs = []
    for split in (1, S):
        C = torch.empty(M, N, dtype=torch.float32, device="cuda:0")
        MLP._gemm(_lib.NFX_MLP_STORE, M, N, K, Ad, K, Bd, K, C=C, split=split, ws=ws)
        outs.append(C.cpu().numpy())
        assert np.max(np.abs(outs[-1] - ref)) < 1e-5 * np.max(np.abs(ref)) + 1e-4, split
    assert not np.array_equal(outs[0], np.zeros_like(outs[0]))
    W0 = RNG.normal(size=(M, N)).astype(np.float32)
    got = []
    for split in (1, S):
        W = torch.tensor(W0, device="cuda:0")
        m, v = torch.zeros_like(W), torch.zeros_like(W)
        H = torch.empty(M, N, dtype=torch.bfloat16, device="cuda:0")
        HT = torch.empty(N, M, dtype=torch.bfloat16, device="cuda:0")
        MLP._gemm(_lib.NFX_MLP_ADAM, M, N, K, Ad, K, Bd, K, C=W, H=H, HT=HT, m=m, v=v, lr_t=1e-2,
                  bc2=1.0 / math.sqrt(1.0 - 0.999), split=split, ws=ws)
        torch.cuda.synchronize()
        assert torch.equal(H.t().contiguous(), HT) and torch.equal(H, W.to(torch.bfloat16))
        got.append(W.cpu().numpy())
    assert np.max(np.abs(got[0] - W0)) > 5e-4  # the step moved the weights (about lr per entry)
    assert np.max(np.abs(got[0] - got[1])) < 1e-6


SMALL = ((6, 128, 128, 128), 256)
# the surrogate's network and training batch: 6 -> 512 x4 (GELU) -> 256, batch 4096
FULL = ((6, 512, 512, 512, 512, 256), 4096)


def make_net(sizes):
    from nofluidx3d_amd.surrogate.mlp import MLP

    return MLP(sizes=sizes, seed=3)


def oracle_params(net):
    Ws = [bf_round(w.cpu().numpy())[:, :n] for w, n in zip(net.W, net.sizes[:-1])]
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b]
    return Ws, bs


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape,weighted", [(SMALL, False), (SMALL, True), (FULL, True)])
def test_forward_and_gradients_match_oracle(ctx, shape, weighted):
    import torch

    sizes, B = shape
    net = make_net(sizes)
    X = RNG.normal(size=(B, 6))
    T = RNG.normal(scale=0.3, size=(B, sizes[-1]))
    w = RNG.uniform(0.1, 4.0, size=B).astype(np.float32) if weighted else None
    net.set_input(X, train=True)
    Y = net.forward_dev(B, train=True)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda:0")
    wd = None if w is None else torch.tensor(w, device="cuda:0")
    gW, gb = net.backward_dev(B, Td, apply_adam=False, row_weight=wd)
    net.sync()
    Ws, bs = oracle_params(net)
    Xr = bf_round(X)
    Yo, _ = OM.forward(Ws, bs, Xr)
    assert rel(Y.cpu().numpy(), Yo) < 3e-2
    loss, gWo, gbo = OM.loss_and_grads(Ws, bs, Xr, T, None if w is None else w.astype(np.float64))
    assert abs(net.loss_dev.item() - loss) < 3e-2 * loss
    for l in range(len(sizes) - 1):
        assert rel(gW[l].cpu().numpy()[:, :Ws[l].shape[1]], gWo[l]) < 3e-2, l
        assert rel(gb[l].cpu().numpy(), gbo[l]) < 3e-2, l
        assert np.all(gW[l].cpu().numpy()[:, Ws[l].shape[1]:] == 0)  # padded input columns


@pytest.mark.parametrize("shape", [SMALL, FULL])
def test_fused_adam_matches_oracle_adam_on_kernel_gradients(ctx, shape):
    import torch

    sizes, B = shape
    a, b = make_net(sizes), make_net(sizes)
    X = RNG.normal(size=(B, 6))
    T = RNG.normal(scale=0.3, size=(B, sizes[-1]))
    Td = torch.tensor(T, dtype=torch.float32, device="cuda:0")
    a.set_input(X, train=True)
    a.forward_dev(B, train=True)
    gW, gb = a.backward_dev(B, Td, apply_adam=False)
    W0 = [w.cpu().numpy().astype(np.float64) for w in b.W]
    b0 = [x.cpu().numpy().astype(np.float64) for x in b.b]
    b.set_input(X, train=True)
    b.train_step(B, Td)
    b.sync()
    G = [g.cpu().numpy().astype(np.float64) for g in gW] + [g.cpu().numpy().astype(np.float64) for g in gb]
    P = W0 + b0
    ms = [np.zeros_like(p) for p in P]
    vs = [np.zeros_like(p) for p in P]
    OM.adam(P, G, ms, vs, 1)
    got = [w.cpu().numpy() for w in b.W] + [x.cpu().numpy() for x in b.b]
    for p, g in zip(P, got):
        assert np.max(np.abs(p - g)) < 1e-6


def test_training_reduces_loss(ctx):
    import torch

    from nofluidx3d_amd.surrogate.mlp import MLP

    net = MLP(sizes=(6, 256, 256, 128), seed=5)
    B = 1024
    X = RNG.uniform(-1, 1, size=(B, 6))
    t = np.linspace(0, 1, 128)
    T = np.sin(2 * X[:, :1] + 3 * t) * (1 + 0.5 * X[:, 1:2]) + X[:, 2:3] * t**2
    Td = torch.tensor(T, dtype=torch.float32, device="cuda:0")
    net.set_input(X, train=True)
    losses = []
    for _ in range(300):
        net.train_step(B, Td)
        net.sync()
        losses.append(net.loss_dev.item())
    assert losses[-1] < 0.05 * losses[0], (losses[0], losses[-1])
