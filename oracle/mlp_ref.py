"""CPU reference of the surrogate MLP's forward and backward passes.

TEST INFRASTRUCTURE ONLY: the GPU kernels (zero-rk_amd/csrc/mlp.hip) are
checked against it; nothing in the product imports it.

The network (no counterpart in the reference) maps features x (BF16) through
dense layers z_l = h_{l-1} W_l + b_l, h_l = gelu(z_l) (identity on the last
layer), with BF16 operands and FP32 accumulation.  The reference rounds the
GEMM operands to BF16 exactly as the kernels do (round to nearest even) and
accumulates in FP64, so the two differ only by FP32 accumulation order.

Gradients of the loss L = mean over the batch and the first n_out outputs of
(out - target)^2 follow the textbook chain rule; gelu is the exact erf form,
gelu'(z) = Phi(z) + z phi(z).
"""
from __future__ import annotations

import math

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = np.isnan(x)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32)
    return np.where(nan, np.float32(np.nan), out).astype(np.float32)


_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    cdf = 0.5 * (1.0 + _erf(z / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)
    return cdf + z * pdf


def _bf16_64(x):
    """BF16 rounding of the kernels' GEMM operands, as FP64 values."""
    return bf16(np.asarray(x).astype(np.float32)).astype(np.float64)


def _exact(x):
    """No rounding: the FP64 network, for derivative checks."""
    return np.asarray(x, dtype=np.float64)


def forward(x_bf16, Ws, bs, rnd=_bf16_64):
    """x: [B, K0] (values already BF16); Ws[l]: [K_l, N_l] FP32 masters (used as
    BF16); returns (out, cache) with out FP64 [B, N_L].  rnd: the rounding of
    GEMM operands (BF16 as the kernels do; _exact for the FP64 network)."""
    h = np.asarray(x_bf16, dtype=np.float64)
    zs, hs = [], [h]
    L = len(Ws)
    for i, (W, b) in enumerate(zip(Ws, bs)):
        z = hs[-1] @ rnd(W) + np.asarray(b, np.float64)
        zs.append(z)
        if i < L - 1:
            hs.append(rnd(gelu(z)))
    return zs[-1], (zs, hs)


def backward(dout, Ws, cache, rnd=_bf16_64):
    """dout: dL/d(out) [B, N_L]; returns (dWs, dbs) in FP64, with the BF16
    rounding of the backward GEMM operands the kernels apply (rnd as in
    forward)."""
    zs, hs = cache
    L = len(Ws)
    dWs, dbs = [None] * L, [None] * L
    dz = np.asarray(dout, np.float64)
    for i in range(L - 1, -1, -1):
        dzb = rnd(dz)
        dWs[i] = hs[i].T @ dzb
        dbs[i] = dz.sum(axis=0)
        if i > 0:
            dh = dzb @ rnd(Ws[i]).T
            dz = dh * gelu_grad(zs[i - 1])
    return dWs, dbs


def mse_grad(out, target, n_out):
    """dL/d(out) of L = mean((out[:, :n_out] - target)^2); zero on padding."""
    B = out.shape[0]
    g = np.zeros_like(out, dtype=np.float64)
    g[:, :n_out] = 2.0 * (out[:, :n_out] - target) / (B * n_out)
    return g
