"""Pins of the FP64 numpy MLP oracle: gradients against torch autograd (float64, CPU) and central
differences; Adam against torch.optim.Adam."""

import numpy as np
import torch

import oracle.mlp as M

RNG = np.random.default_rng(7)


def net(sizes=(6, 24, 24, 10)):
    Ws = [RNG.normal(scale=1 / np.sqrt(a), size=(b, a)) for a, b in zip(sizes[:-1], sizes[1:])]
    bs = [RNG.normal(scale=0.1, size=b) for b in sizes[1:]]
    return Ws, bs


def test_gradients_match_torch_autograd():
    Ws, bs = net()
    X, T = RNG.normal(size=(17, 6)), RNG.normal(size=(17, 10))
    loss, gW, gb = M.loss_and_grads(Ws, bs, X, T)
    tW = [torch.tensor(w, requires_grad=True) for w in Ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    h = torch.tensor(X)
    for l in range(len(tW)):
        h = h @ tW[l].T + tb[l]
        if l < len(tW) - 1:
            h = torch.nn.functional.gelu(h)  # exact (erf) GELU
    tl = ((h - torch.tensor(T)) ** 2).mean()
    tl.backward()
    assert abs(loss - tl.item()) < 1e-14
    for l in range(len(Ws)):
        assert np.allclose(gW[l], tW[l].grad.numpy(), rtol=1e-10, atol=1e-14)
        assert np.allclose(gb[l], tb[l].grad.numpy(), rtol=1e-10, atol=1e-14)


def test_row_weighted_loss_matches_torch_autograd():
    Ws, bs = net()
    X, T = RNG.normal(size=(13, 6)), RNG.normal(size=(13, 10))
    w = RNG.uniform(0.2, 3.0, size=13)
    loss, gW, gb = M.loss_and_grads(Ws, bs, X, T, w)
    tW = [torch.tensor(x, requires_grad=True) for x in Ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    h = torch.tensor(X)
    for l in range(len(tW)):
        h = h @ tW[l].T + tb[l]
        if l < len(tW) - 1:
            h = torch.nn.functional.gelu(h)
    tl = (torch.tensor(w)[:, None] * (h - torch.tensor(T)) ** 2).mean()
    tl.backward()
    assert abs(loss - tl.item()) < 1e-13
    for l in range(len(Ws)):
        assert np.allclose(gW[l], tW[l].grad.numpy(), rtol=1e-10, atol=1e-14)
        assert np.allclose(gb[l], tb[l].grad.numpy(), rtol=1e-10, atol=1e-14)
    # uniform weights reduce to the plain mean
    assert np.isclose(M.loss_and_grads(Ws, bs, X, T, np.full(13, 2.0))[0], 2 * M.loss_and_grads(Ws, bs, X, T)[0])


def test_gradients_match_central_differences():
    Ws, bs = net((6, 8, 5))
    X, T = RNG.normal(size=(9, 6)), RNG.normal(size=(9, 5))
    _, gW, _ = M.loss_and_grads(Ws, bs, X, T)
    h = 1e-6
    for (i, j) in [(0, 0), (3, 5), (7, 2)]:
        Wp = [w.copy() for w in Ws]
        Wm = [w.copy() for w in Ws]
        Wp[0][i, j] += h
        Wm[0][i, j] -= h
        fd = (M.loss_and_grads(Wp, bs, X, T)[0] - M.loss_and_grads(Wm, bs, X, T)[0]) / (2 * h)
        assert abs(fd - gW[0][i, j]) < 1e-8


def test_adam_matches_torch():
    p = RNG.normal(size=(5, 4))
    tp = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m, v = np.zeros_like(p), np.zeros_like(p)
    for t in range(1, 6):
        g = RNG.normal(size=p.shape)
        M.adam([p], [g], [m], [v], t)
        tp.grad = torch.tensor(g)
        opt.step()
    assert np.allclose(p, tp.detach().numpy(), rtol=0, atol=1e-12)
