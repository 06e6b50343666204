"""GPU parity of the analytic Jacobian kernel (zrk_jacobian_device) against the
C oracle's restatement of reactor_nvector_serial.cpp:610-771.

Species block (rows and columns of the species): per entry 1e-12 of
max(|J_ij|, 1e-3 * row max), the reference parity metric for Jacobian
entries, plus the round-off bound of the entry's own summation,
4 k u sum|terms| for an entry fed by k terms +-ROP_s / C_c (u the unit
round-off).  The second part only matters where the terms cancel: a row
whose terms cancel exactly in the oracle (a species on both sides of its
steps) is exactly zero there, so the metric's scale is zero, while rates of
progress that differ in the last bit leave a residue of about one ulp of the
terms on the GPU.  The temperature row d(dT/dt)/dY_j is a sum over all species of
h_i J_ij whose terms cancel (mass conservation makes the column sums of the
species block vanish), so its entries carry the round-off of the largest
term rather than their own: it is compared at 1e-12 of max(|J_ij|,
1e-3 * row max, 1e-6 * column max), the column floor covering that residue.
The temperature column is a forward finite difference (relative step
sqrt(eps)), so it inherits round-off amplified by 1/sqrt(eps): compared at
1e-6 of the column maximum plus four times the oracle's own spread of that
column over one-ulp changes of the state (the same finite difference through
different round-off).
"""
import numpy as np
import pytest

from conftest import load_golden, mech_paths

pytestmark = pytest.mark.gpu

RU = 8.314462618e3


def _term_bound(o, P, Y, T):
    """Per species-block entry: the number of Jacobian terms k and the sum of
    their magnitudes (in the mass-fraction basis) at the floored state, from
    the oracle's rates of progress and the step participant lists."""
    t = o.tables
    nsp = len(Y)
    mw = t.mw
    Yf = np.where(np.abs(Y) < 1e-30, np.where(Y < 0.0, -1e-30, 1e-30), Y)
    rop = o.rates(T, o.conc(T, P, Yf), 1.0e22)[0]
    rho = P / (RU * T) / np.sum(Y / mw)
    t2 = mw / Yf / rho
    S = np.zeros((nsp, nsp))
    K = np.zeros((nsp, nsp))
    for s in range(t.nstep):
        rs = t.reac_sp[t.reac_ptr[s]:t.reac_ptr[s + 1]]
        ps = t.prod_sp[t.prod_ptr[s]:t.prod_ptr[s + 1]]
        for c in rs:
            v = abs(rop[s]) * t2[c]
            for r in list(rs) + list(ps):
                S[r, c] += v
                K[r, c] += 1
    return S * (mw[:, None] / mw[None, :]), K


@pytest.mark.parametrize("name", ["synth_10sp", "synth_53sp", "ref_h2", "ref_h2_no_falloff",
                                  "ref_dme", "synth_160sp"])
def test_jacobian_matches_oracle(torch_gpu, oracle_mod, name):
    from zerork_hip import capi
    torch = torch_gpu
    g = load_golden(name)
    m = capi.Mech(*mech_paths(name))
    o = oracle_mod.Oracle(*mech_paths(name))
    n, nsp = g["Y"].shape
    N = nsp + 1
    print("case", name, n, flush=True)
    dev = torch.device("cuda:0")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    J = torch.zeros(n * N * N, dtype=torch.float64, device=dev)
    m.jacobian(n, t(g["T"]), t(g["P"]), t(g["Y"]), nsp, J)
    torch.cuda.synchronize()
    Jg = J.cpu().numpy().reshape(n, N, N)
    worst_sp = worst_t = 0.0
    for i in range(n):
        Jo = o.jacobian(g["P"][i], np.append(g["Y"][i], g["T"][i]))
        A, B = Jg[i][:nsp, :nsp], Jo[:nsp, :nsp]
        scale = np.maximum(np.abs(B), 1e-3 * np.max(np.abs(B), axis=1, keepdims=True))
        S, K = _term_bound(o, g["P"][i], g["Y"][i], g["T"][i])
        roundoff = 4.0 * K * np.finfo(float).eps * S
        err = np.maximum(np.abs(A - B) - roundoff, 0.0) / (scale + 1e-300)
        worst_sp = max(worst_sp, float(err.max()))
        a, b = Jg[i][nsp, :nsp], Jo[nsp, :nsp]
        tscale = np.maximum(np.maximum(np.abs(b), 1e-3 * np.max(np.abs(b))),
                            1e-6 * np.max(np.abs(Jo[:, :nsp]), axis=0))
        terr = np.abs(a - b) / (tscale + 1e-300)
        worst_t = max(worst_t, float(terr.max()))
        cmax = np.max(np.abs(Jo[:, nsp]))
        y = np.append(g["Y"][i], g["T"][i])
        big = int(np.argmax(g["Y"][i]))
        noise = np.zeros(N)
        for k, to in ((nsp, np.inf), (nsp, -np.inf), (big, np.inf)):
            yp = y.copy()
            yp[k] = np.nextafter(yp[k], to)
            noise = np.maximum(noise, np.abs(o.jacobian(g["P"][i], yp)[:, nsp] - Jo[:, nsp]))
        assert np.all(np.abs(Jg[i][:, nsp] - Jo[:, nsp]) <= 1e-6 * cmax + 4.0 * noise), \
            (i, np.max(np.abs(Jg[i][:, nsp] - Jo[:, nsp]) / cmax), np.max(noise) / cmax)
    print(f"{name}: species block max error {worst_sp:.2e}, temperature row {worst_t:.2e} "
          "(units of the scales above)", flush=True)
    assert worst_sp <= 1e-12 and worst_t <= 1e-12, (worst_sp, worst_t)
