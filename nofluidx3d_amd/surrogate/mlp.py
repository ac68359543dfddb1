"""Fully connected GELU network trained and evaluated with the libnfx MFMA kernels.

Layout (all device tensors):
- layer l: FP32 master weights W[l] [out x in_pad] (PyTorch Linear convention, input padded with
  zero columns to a multiple of 32), FP32 bias b[l]; BF16 copies Wb[l] [out x in_pad] (forward
  operand) and WbT[l] [in_pad x out] (backward-data operand); Adam moments for W and b;
- activations for a batch of B rows: pre-activations Z[l] FP32 [B x out], H[l] BF16 [B x out] and
  its transpose HT[l] [out x B]; gradients dZ[l] BF16 [B x out] and dZT[l] [out x B].

One training step (``train_step``): forward (FWD_GELU epilogues, STORE for the output layer),
MSE loss (optionally row-weighted) and its gradient, then for l = L..1: backward-data GEMM for layer l-1 with the current
BF16 weights (BWD_GELU epilogue), weight-gradient GEMM with the fused Adam update (ADAM epilogue;
split over K = batch into ``split_k`` slices, since its out x in output is too small to fill the
chip), bias gradient + Adam.  Weight initialisation follows torch.nn.Linear (U(-1/sqrt(fan_in),
1/sqrt(fan_in)) for weights and biases) from a seeded generator.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from nofluidx3d_amd import _lib

LR, BETA1, BETA2, EPS = 1e-3, 0.9, 0.999, 1e-8


def _pad32(n):
    return -(-n // 32) * 32


def split_k(M, N, K, min_groups=512, max_split=64):
    """K slices for a GEMM of M x N outputs over K: doubled until the grid has ``min_groups``
    workgroups of 128 x 32 outputs (the weight-gradient GEMMs have few outputs and K = batch),
    each slice a multiple of 32 long"""
    groups = (M // 128) * (N // 32)
    s = 1
    while s < max_split and groups * s < min_groups and K % (64 * s) == 0:
        s *= 2
    return s


class MLP:
    def __init__(self, sizes=(6, 512, 512, 512, 512, 256), seed=0, device="cuda:0"):
        import torch

        from nofluidx3d_amd import openCL

        openCL.initializeOpenCLObjects(int(str(device).split(":")[-1]) if ":" in str(device) else 0)
        self.dev = torch.device(device)
        self.sizes = tuple(sizes)
        self.in_pad = [_pad32(n) for n in sizes[:-1]]
        self.nl = len(sizes) - 1
        g = torch.Generator().manual_seed(seed)
        self.W, self.b = [], []
        for l in range(self.nl):
            fin, fout = sizes[l], sizes[l + 1]
            bound = 1.0 / math.sqrt(fin)
            w = torch.zeros(fout, self.in_pad[l], dtype=torch.float32)
            w[:, :fin] = (torch.rand(fout, fin, generator=g, dtype=torch.float64) * 2 - 1).float() * bound
            bb = ((torch.rand(fout, generator=g, dtype=torch.float64) * 2 - 1) * bound).float()
            self.W.append(w.to(self.dev))
            self.b.append(bb.to(self.dev))
        self._sync_bf16()
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.step = 0
        self._bufs = {}
        self.loss_dev = torch.zeros(1, dtype=torch.float64, device=self.dev)

    def _sync_bf16(self):
        """BF16 weight copies from the FP32 masters (set-up and after loading weights)."""
        import torch

        self.Wb = [w.to(torch.bfloat16).contiguous() for w in self.W]
        self.WbT = [w.t().contiguous().to(torch.bfloat16) for w in self.W]
        torch.cuda.synchronize(self.dev)

    # -------------------------------------------------------------------------------------------
    def buffers(self, B, train):
        import torch

        key = (B, train)
        if key not in self._bufs:
            bf, f32 = torch.bfloat16, torch.float32
            d = {"X": torch.zeros(B, self.in_pad[0], dtype=bf, device=self.dev),
                 "XT": torch.zeros(self.in_pad[0], B, dtype=bf, device=self.dev) if train else None,
                 "H": [torch.empty(B, n, dtype=bf, device=self.dev) for n in self.sizes[1:-1]],
                 "Y": torch.empty(B, self.sizes[-1], dtype=f32, device=self.dev)}
            if train:
                # weight-gradient GEMMs (M = out, N = in_pad, K = B) in K slices, one shared workspace
                d["split"] = [split_k(n_out, n_in, B) for n_out, n_in in zip(self.sizes[1:], self.in_pad)]
                ws = max(s * n_out * n_in for s, n_out, n_in in zip(d["split"], self.sizes[1:], self.in_pad))
                d["ws"] = torch.empty(ws, dtype=f32, device=self.dev)
                d["Z"] = [torch.empty(B, n, dtype=f32, device=self.dev) for n in self.sizes[1:-1]]
                d["HT"] = [torch.empty(n, B, dtype=bf, device=self.dev) for n in self.sizes[1:-1]]
                d["dZ"] = [torch.empty(B, n, dtype=bf, device=self.dev) for n in self.sizes[1:]]
                d["dZT"] = [torch.empty(n, B, dtype=bf, device=self.dev) for n in self.sizes[1:]]
            self._bufs[key] = d
        return self._bufs[key]

    @staticmethod
    def _gemm(epi, M, N, K, A, lda, B, ldb, bias=None, C=None, Z=None, H=None, HT=None, m=None, v=None, lr_t=0.0,
              bc2=1.0, grad_scale=1.0, split=1, ws=None):
        from nofluidx3d_amd import openCL

        p = lambda t: None if t is None else t.data_ptr()
        g = _lib.MlpGemm(epi, M, N, K, p(A), lda, p(B), ldb, p(bias), p(C), p(Z), p(H), p(HT), p(m), p(v),
                         lr_t, bc2, BETA1, BETA2, EPS, grad_scale, p(ws) if split > 1 else None, split)
        _lib.call("nfx_mlp_gemm", openCL.handle(), ctypes.byref(g))

    def set_input(self, X, train=False):
        """X: (B, n_in) float features -> BF16 padded device input (and its transpose for training)."""
        import torch

        B = X.shape[0]
        bufs = self.buffers(B, train)
        Xp = torch.zeros(B, self.in_pad[0], dtype=torch.float32)
        Xp[:, :self.sizes[0]] = torch.as_tensor(np.asarray(X, np.float32))
        bufs["X"].copy_(Xp.to(torch.bfloat16))
        if train:
            bufs["XT"].copy_(Xp.t().contiguous().to(torch.bfloat16))
        torch.cuda.synchronize(self.dev)
        return bufs

    def forward_dev(self, B, train=False):
        bufs = self.buffers(B, train)
        inp = bufs["X"]
        for l in range(self.nl):
            K, N = self.in_pad[l], self.sizes[l + 1]
            if l < self.nl - 1:
                self._gemm(_lib.NFX_MLP_FWD_GELU, B, N, K, inp, K, self.Wb[l], K, bias=self.b[l],
                           C=bufs["Z"][l] if train else None, H=bufs["H"][l], HT=bufs["HT"][l] if train else None)
                inp = bufs["H"][l]
            else:
                self._gemm(_lib.NFX_MLP_STORE, B, N, K, inp, K, self.Wb[l], K, bias=self.b[l], C=bufs["Y"])
        return bufs["Y"]

    def predict(self, X, chunk=65536):
        """Network output (B, n_out) float32 numpy for features X (rows padded to a multiple of 128)."""
        X = np.asarray(X, np.float32)
        out = np.empty((len(X), self.sizes[-1]), np.float32)
        for s in range(0, len(X), chunk):
            xb = X[s:s + chunk]
            B = -(-len(xb) // 128) * 128
            xp = np.zeros((B, X.shape[1]), np.float32)
            xp[:len(xb)] = xb
            self.set_input(xp)
            y = self.forward_dev(B)
            self.sync()
            out[s:s + len(xb)] = y[:len(xb)].cpu().numpy()
        return out

    def sync(self):
        import torch

        from nofluidx3d_amd import openCL

        _lib.call("nfx_stream_sync", openCL.handle())
        torch.cuda.synchronize(self.dev)

    def backward_dev(self, B, target, apply_adam=True, row_weight=None):
        """MSE gradient of the last forward pass (rows weighted by the FP32 device tensor
        ``row_weight`` when given); Adam updates (apply_adam) or raw gradients (returned as lists
        of FP32 tensors dW, db)."""
        import torch

        from nofluidx3d_amd import openCL

        bufs = self.buffers(B, True)
        L = self.nl
        self.loss_dev.zero_()
        _lib.call("nfx_mlp_mse", openCL.handle(), bufs["Y"].data_ptr(), target.data_ptr(),
                  None if row_weight is None else row_weight.data_ptr(), B, self.sizes[-1],
                  bufs["dZ"][L - 1].data_ptr(), bufs["dZT"][L - 1].data_ptr(), self.loss_dev.data_ptr())
        if apply_adam:
            self.step += 1
            t = self.step
            lr_t = LR / (1 - BETA1**t)
            bc2 = 1.0 / math.sqrt(1 - BETA2**t)
        grads_w, grads_b = [None] * L, [None] * L
        for l in range(L - 1, -1, -1):
            K_in, N_out = self.in_pad[l], self.sizes[l + 1]
            prev_HT = bufs["XT"] if l == 0 else bufs["HT"][l - 1]
            if l > 0:  # gradient into layer l-1 with the current weights, before they change
                self._gemm(_lib.NFX_MLP_BWD_GELU, B, K_in, N_out, bufs["dZ"][l], N_out, self.WbT[l], N_out,
                           Z=bufs["Z"][l - 1], H=bufs["dZ"][l - 1], HT=bufs["dZT"][l - 1])
            S = bufs["split"][l]
            if apply_adam:
                self._gemm(_lib.NFX_MLP_ADAM, N_out, K_in, B, bufs["dZT"][l], B, prev_HT, B, C=self.W[l],
                           H=self.Wb[l], HT=self.WbT[l], m=self.mW[l], v=self.vW[l], lr_t=lr_t, bc2=bc2, split=S,
                           ws=bufs["ws"])
                _lib.call("nfx_mlp_bias_adam", openCL.handle(), bufs["dZT"][l].data_ptr(), B, N_out,
                          self.b[l].data_ptr(), self.mb[l].data_ptr(), self.vb[l].data_ptr(), lr_t, bc2, BETA1,
                          BETA2, EPS, 1.0)
            else:
                gw = torch.empty(N_out, K_in, dtype=torch.float32, device=self.dev)
                self._gemm(_lib.NFX_MLP_STORE, N_out, K_in, B, bufs["dZT"][l], B, prev_HT, B, C=gw, split=S,
                           ws=bufs["ws"])
                grads_w[l] = gw
                grads_b[l] = bufs["dZT"][l].float().sum(1)
        return grads_w, grads_b

    def train_step(self, B, target, row_weight=None):
        """One Adam step on the batch already set with ``set_input(X, train=True)``; returns the
        device loss tensor (FP64, row-weighted mean squared error of the forward pass)."""
        self.forward_dev(B, train=True)
        self.backward_dev(B, target, apply_adam=True, row_weight=row_weight)
        return self.loss_dev

    # -------------------------------------------------------------------------------------------
    def state(self):
        return {f"W{l}": self.W[l].cpu().numpy() for l in range(self.nl)} | {
            f"b{l}": self.b[l].cpu().numpy() for l in range(self.nl)}

    def load_state(self, st):
        import torch

        for l in range(self.nl):
            self.W[l].copy_(torch.as_tensor(st[f"W{l}"]))
            self.b[l].copy_(torch.as_tensor(st[f"b{l}"]))
        self._sync_bf16()
