"""Pins the surrogate MLP oracle (oracle/mlp_ref.py), which the GPU kernels of
zero-rk_amd/csrc/mlp.hip are compared with, by what calculus fixes rather than
by a restatement of its own formulas:

* with the BF16 rounding switched off (the FP64 network), every gradient
  entry of backward() equals the central finite difference of the loss
  L = mean((out - target)^2) over the first n_out outputs, for the weights of
  every layer, the biases and the network's 3 x 256 GELU shape as well as a
  small one;
* gelu_grad() equals the derivative of gelu() (central differences), and
  gelu() meets its closed-form values (0 at 0, z for large z, 0 for large -z,
  z Phi(z) at z = 1);
* mistakes planted in a copy of the oracle (a dropped z phi(z) term of
  GELU', a weight matrix not transposed in the backward product, the bias
  gradient of the wrong layer) each fail the same finite-difference check,
  so the check can tell them apart.
"""
import math

import numpy as np
import pytest

from oracle import mlp_ref


def _net(sizes, seed):
    rng = np.random.default_rng(seed)
    Ws = [rng.standard_normal((a, b)) / math.sqrt(a) for a, b in zip(sizes[:-1], sizes[1:])]
    bs = [0.1 * rng.standard_normal(b) for b in sizes[1:]]
    return Ws, bs


def _loss(x, Ws, bs, target, n_out):
    out, _ = mlp_ref.forward(x, Ws, bs, rnd=mlp_ref._exact)
    return float(np.mean((out[:, :n_out] - target) ** 2))


def _fd_errors(sizes, batch, n_out, seed, backward=mlp_ref.backward, samples=12):
    """max over sampled entries of |analytic - central FD| / (|FD| + scale)"""
    rng = np.random.default_rng(seed + 1)
    Ws, bs = _net(sizes, seed)
    x = rng.standard_normal((batch, sizes[0]))
    target = rng.standard_normal((batch, n_out))
    out, cache = mlp_ref.forward(x, Ws, bs, rnd=mlp_ref._exact)
    dWs, dbs = backward(mlp_ref.mse_grad(out, target, n_out), Ws, cache, rnd=mlp_ref._exact)
    worst = 0.0
    for layer in range(len(Ws)):
        params = [(Ws, dWs, (int(rng.integers(Ws[layer].shape[0])),
                             int(rng.integers(Ws[layer].shape[1])))) for _ in range(samples)]
        params += [(bs, dbs, (int(rng.integers(bs[layer].shape[0])),)) for _ in range(samples // 2)]
        for P, G, idx in params:
            v0 = P[layer][idx]
            eps = 1e-5 * max(1.0, abs(v0))
            P[layer][idx] = v0 + eps
            lp = _loss(x, Ws, bs, target, n_out)
            P[layer][idx] = v0 - eps
            lm = _loss(x, Ws, bs, target, n_out)
            P[layer][idx] = v0
            fd = (lp - lm) / (2.0 * eps)
            scale = 1e-6 * max(abs(g) for g in G[layer].ravel()[:64]) + 1e-12
            worst = max(worst, abs(G[layer][idx] - fd) / (abs(fd) + scale))
    return worst


@pytest.mark.parametrize("sizes,n_out", [([13, 48, 48, 11], 10), ([56, 256, 256, 256, 64], 54)])
def test_backward_matches_finite_differences(sizes, n_out):
    assert _fd_errors(sizes, batch=6, n_out=n_out, seed=3) < 1e-6


def test_gelu_and_its_derivative():
    z = np.linspace(-6.0, 6.0, 241)
    h = 1e-6
    fd = (mlp_ref.gelu(z + h) - mlp_ref.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp_ref.gelu_grad(z), fd, rtol=1e-8, atol=1e-9)
    assert mlp_ref.gelu(np.array([0.0]))[0] == 0.0
    np.testing.assert_allclose(mlp_ref.gelu(np.array([30.0, -30.0])), [30.0, 0.0], atol=1e-12)
    phi1 = 0.5 * (1.0 + math.erf(1.0 / math.sqrt(2.0)))
    np.testing.assert_allclose(mlp_ref.gelu(np.array([1.0]))[0], phi1, rtol=1e-15)


def _planted(kind):
    """backward() with one planted mistake (a copy: the oracle is unchanged)"""
    def bw(dout, Ws, cache, rnd=mlp_ref._exact):
        zs, hs = cache
        L = len(Ws)
        dWs, dbs = [None] * L, [None] * L
        dz = np.asarray(dout, np.float64)
        for i in range(L - 1, -1, -1):
            dWs[i] = hs[i].T @ dz
            dbs[i] = dz.sum(axis=0)
            if i > 0:
                # square layers let the untransposed product run
                dh = dz @ (Ws[i] if kind == "transpose" else Ws[i].T)
                z = zs[i - 1]
                if kind == "gelu":  # Phi(z) without the z phi(z) term
                    dz = dh * 0.5 * (1.0 + np.vectorize(math.erf)(z / math.sqrt(2.0)))
                else:
                    dz = dh * mlp_ref.gelu_grad(z)
        if kind == "bias":
            dbs[0], dbs[1] = dbs[1], dbs[0]
        return dWs, dbs
    return bw


@pytest.mark.parametrize("kind", ["gelu", "transpose", "bias"])
def test_planted_mistakes_fail_the_check(kind):
    sizes = [24, 24, 24, 24]  # square hidden layers, so each mistake still runs
    assert _fd_errors(sizes, batch=6, n_out=20, seed=5, backward=_planted(kind)) > 1e-3
