"""Learned surrogate of one reactor step (no counterpart in the reference).

An MLP maps the features [ln T, ln P, ln max(Y_j, 1e-20), ln dt] of a reactor
to its end-of-step changes in log space, ln(max(Y1_j, 1e-20)/max(Y0_j, 1e-20))
and ln(T1/T0), normalised per output (relative changes keep minor species,
which span many decades, as well conditioned as major ones).  Forward,
backward and Adam run in the library's own BF16-input / FP32-accumulate
MFMA kernels (include/zrk_surrogate.h); PyTorch only holds the device arrays.

`solve()` is the surrogate-accelerated step: the network proposes each
reactor's state at the nodes t_k = tau_k dt of its step (one forward pass per
node, dt replaced by t_k in the features), an FP64 acceptance test keeps the
trajectories that pass (backward Euler on the first interval, the
trapezoidal rule on the others, each solved by at most a few Newton
iterations started from the proposal, and a local-error estimate within the
integrator's tolerance; zrk_surrogate_accept), and the plain BDF integrates
the others with their own dt.  The network-free proposal "hold" (every node
at the initial state) goes through the same test, which measures what the
network gains.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import capi


def _lib():
    L = capi.lib()
    if not getattr(L, "_surrogate_declared", False):
        import ctypes as C
        P, I, D, F, Lg = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_long
        L.zrk_integrate_device_select.argtypes = [P, C.POINTER(capi.BdfOptions), I, P, P, P, P, P,
                                                  I, P, P, P]
        L.zrk_mlp_features.argtypes = [P, I, P, P, P, I, P, D, P, P, P, I, P]
        L.zrk_gemm_bf16.argtypes = [I, I, I, P, Lg, Lg, P, Lg, Lg, P, Lg, Lg, I, P]
        L.zrk_mlp_bias_act.argtypes = [I, I, P, P, P, P, P, I, P]
        L.zrk_mlp_act_backward.argtypes = [I, I, P, P, P, P, I, P]
        L.zrk_mlp_col_sum.argtypes = [I, I, P, P, P]
        L.zrk_mlp_adam.argtypes = [Lg, P, P, P, P, P, F, F, F, F, I, P]
        L.zrk_to_bf16.argtypes = [Lg, P, P, P]
        L.zrk_surrogate_accept.argtypes = [P, C.POINTER(capi.BdfOptions), I, P, D, P, P, P, I, P,
                                           P, P, I, I, P, P, P]
        for f in ("zrk_integrate_device_select", "zrk_mlp_features", "zrk_gemm_bf16",
                  "zrk_mlp_bias_act", "zrk_mlp_act_backward", "zrk_mlp_col_sum", "zrk_mlp_adam",
                  "zrk_to_bf16", "zrk_surrogate_accept"):
            getattr(L, f).restype = C.c_int
        L._surrogate_declared = True
    return L


def _p(t):
    return None if t is None else t.data_ptr()


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def gemm(A, sA, B, sB, C, M, N, K, accumulate=False, stream=0):
    """C[M,N] (f32) (+)= A[M,K] . B[K,N]; A/B bf16 tensors with element strides."""
    _chk(_lib().zrk_gemm_bf16(M, N, K, _p(A), sA[0], sA[1], _p(B), sB[0], sB[1], _p(C), N, 1,
                              int(accumulate), stream), "zrk_gemm_bf16")


@dataclass
class Net:
    """Dense layers with FP32 masters and BF16 copies on the device."""
    sizes: list
    W: list = field(default_factory=list)     # f32 [K, N]
    b: list = field(default_factory=list)     # f32 [N]
    Wb: list = field(default_factory=list)    # bf16 [K, N]

    @staticmethod
    def create(torch, sizes, seed, device):
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = Net(list(sizes))
        for k, m in zip(sizes[:-1], sizes[1:]):
            w = torch.randn(k, m, generator=g, dtype=torch.float32) * math.sqrt(2.0 / k)
            n.W.append(w.to(device))
            n.b.append(torch.zeros(m, dtype=torch.float32, device=device))
            n.Wb.append(torch.empty(k, m, dtype=torch.bfloat16, device=device))
        n.refresh()
        return n

    def refresh(self):
        for w, wb in zip(self.W, self.Wb):
            _chk(_lib().zrk_to_bf16(w.numel(), _p(w), _p(wb), 0), "zrk_to_bf16")


class Surrogate:
    """Training and inference of the step surrogate for one mechanism
    (nsp <= 63; constant pressure)."""

    def __init__(self, torch, mech: capi.Mech, hidden=256, layers=3, seed=0, device="cuda:0"):
        self.torch = torch
        self.mech = mech
        self.nsp = mech.nsp
        self.nfeat = self.nsp + 3
        self.kin = ((self.nfeat + 15) // 16) * 16
        self.nout = self.nsp + 1
        self.kout = ((self.nout + 15) // 16) * 16
        self.device = device
        sizes = [self.kin] + [hidden] * layers + [self.kout]
        self.net = Net.create(torch, sizes, seed, device)
        self.mu = torch.zeros(self.kin, dtype=torch.float32, device=device)
        self.sig = torch.ones(self.kin, dtype=torch.float32, device=device)
        self.tmu = np.zeros(self.nout)
        self.tsig = np.ones(self.nout)
        self.adam_m = [torch.zeros_like(w) for w in self.net.W]
        self.adam_v = [torch.zeros_like(w) for w in self.net.W]
        self.adam_mb = [torch.zeros_like(b) for b in self.net.b]
        self.adam_vb = [torch.zeros_like(b) for b in self.net.b]
        self.step = 0

    # -- features and normalisation -------------------------------------------
    FLOOR = 1e-20

    @classmethod
    def raw_targets(cls, T0, Y0, T1, Y1):
        """Log-space changes of a step (host arrays): [n, nsp + 1]."""
        f = cls.FLOOR
        return np.column_stack([np.log(np.maximum(Y1, f)) - np.log(np.maximum(Y0, f)),
                                np.log(T1 / T0)])

    def fit_normalisation(self, T, P, Y, dt, T1, Y1):
        """Feature and target statistics from a training set (host arrays):
        initial states (T, P, Y), step sizes dt and end states (T1, Y1)."""
        x = np.column_stack([np.log(T), np.log(P), np.log(np.maximum(Y, self.FLOOR)), np.log(dt)])
        mu = np.zeros(self.kin, np.float32)
        sig = np.ones(self.kin, np.float32)
        mu[:self.nfeat] = x.mean(axis=0)
        sig[:self.nfeat] = np.maximum(x.std(axis=0), 1e-6)
        self.mu.copy_(self.torch.as_tensor(mu))
        self.sig.copy_(self.torch.as_tensor(sig))
        tgt = self.raw_targets(T, Y, T1, Y1)
        self.tmu = tgt.mean(axis=0)
        self.tsig = np.maximum(tgt.std(axis=0), 1e-12)

    def target_matrix(self, T0, Y0, T1, Y1):
        """Normalised, zero-padded f32 training targets [n, kout] (host)."""
        tgt = np.zeros((len(T0), self.kout), np.float32)
        tgt[:, :self.nout] = (self.raw_targets(T0, Y0, T1, Y1) - self.tmu) / self.tsig
        return tgt

    def features(self, T, P, Y, dt, stream=0):
        n = T.numel()
        X = self.torch.empty((n, self.kin), dtype=self.torch.bfloat16, device=self.device)
        _chk(_lib().zrk_mlp_features(self.mech.h, n, _p(T), _p(P), _p(Y), self.nsp, _p(dt), 0.0,
                                     _p(self.mu), _p(self.sig), _p(X), self.kin, stream),
             "zrk_mlp_features")
        return X

    # -- network ----------------------------------------------------------------
    def forward(self, X, cache=False, stream=0):
        """X: bf16 [B, kin] -> f32 [B, kout] (normalised outputs)."""
        t = self.torch
        B = X.shape[0]
        h = X
        zs, hs = [], [X]
        L = len(self.net.W)
        out = None
        for i in range(L):
            K, N = self.net.sizes[i], self.net.sizes[i + 1]
            C = t.empty((B, N), dtype=t.float32, device=self.device)
            gemm(h, (K, 1), self.net.Wb[i], (N, 1), C, B, N, K, stream=stream)
            last = i == L - 1
            Z = t.empty_like(C) if (cache or last) else None
            Yb = None if last else t.empty((B, N), dtype=t.bfloat16, device=self.device)
            _chk(_lib().zrk_mlp_bias_act(B, N, _p(C), _p(self.net.b[i]), _p(Z), _p(Yb), None,
                                         0 if last else 1, stream), "zrk_mlp_bias_act")
            if last:
                out = Z
            else:
                if cache:
                    zs.append(Z)
                    hs.append(Yb)
                h = Yb
        return (out, (zs, hs)) if cache else out

    def backward(self, dout, cache, stream=0):
        """dout: f32 [B, kout] = dL/d(out); returns (dW, db) lists (f32)."""
        t = self.torch
        zs, hs = cache
        L = len(self.net.W)
        B = dout.shape[0]
        dW, db = [None] * L, [None] * L
        dz = t.empty_like(dout)
        dzb = t.empty(dout.shape, dtype=t.bfloat16, device=self.device)
        _chk(_lib().zrk_mlp_act_backward(B, dout.shape[1], _p(dout), None, _p(dz), _p(dzb), 0,
                                         stream), "zrk_mlp_act_backward")
        for i in range(L - 1, -1, -1):
            K, N = self.net.sizes[i], self.net.sizes[i + 1]
            dW[i] = t.empty((K, N), dtype=t.float32, device=self.device)
            # dW = h^T dz: A[k_row, b] = h[b, k_row]
            gemm(hs[i], (1, K), dzb, (N, 1), dW[i], K, N, B, stream=stream)
            db[i] = t.empty(N, dtype=t.float32, device=self.device)
            _chk(_lib().zrk_mlp_col_sum(B, N, _p(dz), _p(db[i]), stream), "zrk_mlp_col_sum")
            if i > 0:
                # dh = dz W^T: B[n, k] = W[k, n]
                dh = t.empty((B, K), dtype=t.float32, device=self.device)
                gemm(dzb, (N, 1), self.net.Wb[i], (1, N), dh, B, K, N, stream=stream)
                dz = t.empty_like(dh)
                dzb = t.empty(dh.shape, dtype=t.bfloat16, device=self.device)
                _chk(_lib().zrk_mlp_act_backward(B, K, _p(dh), _p(zs[i - 1]), _p(dz), _p(dzb), 1,
                                                 stream), "zrk_mlp_act_backward")
        return dW, db

    def adam(self, dW, db, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=0):
        self.step += 1
        for i in range(len(self.net.W)):
            _chk(_lib().zrk_mlp_adam(self.net.W[i].numel(), _p(self.net.W[i]), _p(dW[i]),
                                     _p(self.adam_m[i]), _p(self.adam_v[i]), _p(self.net.Wb[i]),
                                     lr, b1, b2, eps, self.step, stream), "zrk_mlp_adam")
            _chk(_lib().zrk_mlp_adam(self.net.b[i].numel(), _p(self.net.b[i]), _p(db[i]),
                                     _p(self.adam_mb[i]), _p(self.adam_vb[i]), None, lr, b1, b2,
                                     eps, self.step, stream), "zrk_mlp_adam")

    def train(self, X, target, epochs=10, batch=4096, lr=1e-3, seed=0, log=None):
        """X: bf16 [n, kin]; target: f32 [n, kout] (normalised, zero padded)."""
        t = self.torch
        n = X.shape[0]
        g = t.Generator(device="cpu").manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = t.randperm(n, generator=g).to(self.device)
            tot = 0.0
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                xb, tb = X[idx].contiguous(), target[idx].contiguous()
                out, cache = self.forward(xb, cache=True)
                diff = out[:, :self.nout] - tb[:, :self.nout]
                dout = t.zeros_like(out)
                dout[:, :self.nout] = diff * (2.0 / (batch * self.nout))
                dW, db = self.backward(dout, cache)
                self.adam(dW, db, lr=lr)
                tot += float((diff * diff).mean()) * batch
            losses.append(tot / max(1, (n // batch) * batch))
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.4e}")
        return losses

    # -- proposals and the accelerated step -------------------------------------
    # node fractions of the step: the first interval is short (its backward
    # Euler step absorbs the fast initial transient), the others grow
    TAU = (1.0e-3, 4.0e-3, 0.015, 0.05, 0.13, 0.3, 0.6, 1.0)

    def propose(self, T, P, Y, dt):
        """Proposed states (T1, Y1) (device f64) after per-reactor times dt."""
        t = self.torch
        out = self.forward(self.features(T, P, Y, dt))
        z = out[:, :self.nout].double() * t.as_tensor(self.tsig, device=self.device) + \
            t.as_tensor(self.tmu, device=self.device)
        Y1 = t.exp(t.log(Y.clamp_min(self.FLOOR)) + z[:, :self.nsp])
        Y1 = Y1 / Y1.sum(dim=1, keepdim=True)
        T1 = T * t.exp(z[:, self.nsp])
        return T1.contiguous(), Y1.contiguous()

    def propose_trajectory(self, T, P, Y, dt, tau=TAU):
        """The proposal at every node t_k = tau_k dt: (Tp [m n], Yp [m n, nsp])."""
        t = self.torch
        Ts, Ys = [], []
        for f in tau:
            T1, Y1 = self.propose(T, P, Y, dt * f)
            Ts.append(T1)
            Ys.append(Y1)
        return t.cat(Ts).contiguous(), t.cat(Ys).contiguous()

    def accept(self, T0, P, Y0, Tp, Yp, dt, opts, tau=TAU, max_newton=3):
        """FP64 acceptance of proposed trajectories (zrk_surrogate_accept);
        returns (accept int32 [n], res [n, 4]) and writes an accepted
        reactor's converged end state over its last node."""
        t = self.torch
        n = T0.numel()
        acc = t.empty(n, dtype=t.int32, device=self.device)
        res = t.empty(4 * n, dtype=t.float64, device=self.device)
        tv = t.as_tensor(tau, dtype=t.float64, device=self.device)
        import ctypes as C
        _chk(_lib().zrk_surrogate_accept(self.mech.h, C.byref(opts), n, _p(dt), 0.0, _p(T0), _p(P),
                                         _p(Y0), len(tau), _p(tv), _p(Tp), _p(Yp), self.nsp,
                                         max_newton, _p(acc), _p(res), 0),
             "zrk_surrogate_accept")
        return acc, res.view(n, 4)

    def solve(self, T, P, Y, dt, opts, proposal="mlp", tau=TAU, max_newton=3):
        """Surrogate-accelerated step, in place on T and Y.  proposal: "mlp" or
        "hold" (every node at the initial state, the same test without the
        network).  Returns the accepted mask (device int32), the BDF step
        counts of the rest and the acceptance test's per-reactor results
        (res [n, 4], see accept)."""
        t = self.torch
        n = T.numel()
        m = len(tau)
        if proposal == "mlp":
            Tp, Yp = self.propose_trajectory(T, P, Y, dt, tau)
        else:
            Tp, Yp = T.repeat(m).contiguous(), Y.repeat(m, 1).contiguous()
        acc, res = self.accept(T, P, Y, Tp, Yp, dt, opts, tau, max_newton)
        rej = t.nonzero(acc == 0).flatten().to(t.int32).contiguous()
        keep = acc.bool()
        T[keep] = Tp[(m - 1) * n:][keep]
        Y[keep] = Yp[(m - 1) * n:][keep]
        steps = t.zeros(n, dtype=t.int32, device=self.device)
        if rej.numel():
            import ctypes as C
            _chk(_lib().zrk_integrate_device_select(self.mech.h, C.byref(opts), rej.numel(), _p(rej),
                                                    _p(dt), _p(T), _p(P), _p(Y), self.nsp,
                                                    _p(steps), None, 0),
                 "zrk_integrate_device_select")
        return acc, steps, res
