"""Algorithmic work of the chemistry kernels (HBM bytes and FP64 operations),
computed from a loaded mechanism (capi.Mech).

Counts follow the formulas the kernels evaluate (rate_const.cpp:539-1012,
perf_net.cpp:190-275, reactor_constant_pressure_cpu.cpp:68-133 and the
Newton linear algebra).  Operation convention: add, sub, mul, div = 1;
fused multiply-add = 2; each exp/log/log10/pow = 1 (their internal
polynomial work is not counted, so the FLOP rates derived from these
counts are lower bounds).
"""
from __future__ import annotations

import numpy as np


def rate_bytes_per_reactor(nsp: int, nstep: int) -> int:
    """HBM bytes of one RHS evaluation by the standalone rate kernel: read the
    state (T, P, Y: nsp+2 doubles), write dY/dt, dT/dt (nsp+1) and the per-step
    rates of progress (nstep) that the Jacobian reuses."""
    return 8 * ((nsp + 2) + (nsp + 1) + nstep)


class Work:
    def __init__(self, mech):
        s = mech.steps()
        r = mech.reactions()
        self.nsp, self.nstep = mech.nsp, mech.nstep
        self.reac_n = np.diff(s["reac_ptr"])
        self.prod_n = np.diff(s["prod_ptr"])
        self.r = r

    def rhs_ops(self) -> int:
        nsp = self.nsp
        ops = 12                          # T powers, 1/T, ln T, ln(Patm/RuT)
        ops += nsp * (2 + 9 + 12 + 4)     # Y/MW, cp poly, h and g polys, C_j, cp_j
        ops += 3 * nsp                    # mass, cp and Csum reductions
        r = self.r
        for j in range(len(r["fwd_step"])):
            f = int(r["fwd_step"][j])
            ops += 5 + 3 + int(self.reac_n[f])    # Arrhenius + exp, limiter, x C
            if r["rev_type"][j] == 1:
                ops += 5 + 3 + int(self.prod_n[f])
            elif r["rev_type"][j] == 2:
                ops += int(self.reac_n[f] + self.prod_n[f]) + 4 + 3 + int(self.prod_n[f])
            ne = int(r["n_enhanced"][j])
            if ne >= 0:
                ops += 2 * ne + 3
            ft = int(r["falloff_type"][j])
            if ft >= 0:
                ops += 5 + 4 + (30 if ft > 0 else 0)  # klow, Pr, Troe broadening
        ops += int(self.reac_n.sum() + self.prod_n.sum())  # creation/destruction gathers
        ops += nsp * 6 + 6                # net, dY/dt, energy sum, dT/dt
        return ops

    def jacobian_ops(self) -> int:
        terms = int(np.sum(self.reac_n * (self.reac_n + self.prod_n)))
        nsp = self.nsp
        return 2 * self.rhs_ops() + 2 * terms + nsp * (2 * nsp + 8) + 3 * (nsp + 1)

    @staticmethod
    def inverse_ops(n: int) -> int:
        """Gauss-Jordan inverse of the n x n Newton matrix."""
        return 2 * n ** 3 + 2 * n * n

    @staticmethod
    def solve_ops(n: int) -> int:
        """Newton solve with the explicit inverse (matrix-vector product)."""
        return 2 * n * n

    def bdf_ops(self, nfe, nje, nsetups, nni) -> int:
        n = self.nsp + 1
        return int(nfe * self.rhs_ops() + nje * self.jacobian_ops() + nsetups * self.inverse_ops(n)
                   + nni * (self.solve_ops(n) + 8 * n))
