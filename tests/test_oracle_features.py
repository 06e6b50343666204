"""Pins the oracle's SRI, PLOG and non-integer (fractional stoichiometry,
FORD/RORD) reactions with the closed forms of the reference's own tests, beside
the golden vectors of its rate core (test_oracle_golden.py):

* SRI 3- and 5-parameter blending: tests/reactions/sri_gtest.cpp:255-300 over
  its temperature/pressure grid;
* PLOG: tests/reactions/plog_gtest.cpp (single pressure at, under and over the
  table, negative and multiple A-factors, reverse from the equilibrium
  constant, log-p interpolation over 3 and 5 pressures, split Arrhenius lines);
* FORD: tests/reactions/ford_propane_gtest.cpp:360-440 (rate coefficients with
  the real-order unit conversion, rates of progress, species rates);
* non-integer Jacobian terms: against central differences of the frozen-density
  RHS on mechanisms whose reactions are all non-integer and have no third
  bodies, where the reference formula (+-nu p ROP / C) is the exact derivative.
"""
import numpy as np
import pytest

from conftest import mech_paths

RU = 8.314462618e3
P_ATM = 1.01325e5
E2T = 4184.0 / RU  # cal/mol -> K


def _uniform_conc(o, T, P):
    n = o.nsp
    return np.full(n, 1.0 / n) * P / (RU * T)


def _k_by_reaction(o, T, C):
    K = o.rate_constants(T, C)
    t = o.tables
    kf = np.zeros(t.nrxn)
    kr = np.zeros(t.nrxn)
    for s, (rx, d) in enumerate(zip(t.step_rxn, t.step_dir)):
        (kf if d == 1 else kr)[rx] = K[s]
    return kf, kr


def _near(a, b, tol):
    w = 0.5 * (abs(a) + abs(b))
    return w < tol * tol or abs(a - b) / w <= tol


def test_sri_closed_form(oracle_mod):
    o = oracle_mod.Oracle(*mech_paths("ref_sri"))

    def kfwd(T, P, d, e):
        M = 0.001 * P / (RU * T)
        k_low = 8.0e26 * T ** -3.0
        k_inf = 6.0e16 * T ** -1.0
        Pr = k_low * M / k_inf
        x = 1.0 / (1.0 + np.log10(Pr) ** 2)
        F = d * (0.45 * np.exp(-797.0 / T) + np.exp(-T / 979.0)) ** x * T ** e
        return k_inf * 0.001 * (Pr / (1.0 + Pr)) * F

    T = 300.0
    for _ in range(36):
        P = 1.0
        for _ in range(24):
            kf, _ = _k_by_reaction(o, T, _uniform_conc(o, T, P))
            assert _near(kf[0], kfwd(T, P, 1.0, 0.0), 1e-13), (T, P)
            assert _near(kf[1], kfwd(T, P, 0.9, 2.0), 1e-13), (T, P)
            P *= 2.0
        T += 100.0


def _keq_h2(T):
    # A + A <=> A2 with the constant-cp thermo (plog_gtest.cpp:565-581)
    s1 = 2.5 * np.log(T) - 4.46703359e-01
    s2 = 3.5 * np.log(T) - 4.22439182e+00
    h1 = 2.5 + 2.54737667e+04 / T
    h2 = 3.5 - 1.04352500e+03 / T
    return (P_ATM / (RU * T)) ** -1.0 * np.exp((-2 * s1 + s2) - (-2 * h1 + h2))


def test_plog_closed_forms(oracle_mod):
    o = oracle_mod.Oracle(*mech_paths("ref_plog"))
    tol = 1e-13
    base = lambda T: 1.0e6 * T ** 2.0 * np.exp(-1000.0 / T)
    T = 1000.0
    for P in (P_ATM, 0.1 * P_ATM, 10.0 * P_ATM):  # at, under and over the table
        kf, kr = _k_by_reaction(o, T, _uniform_conc(o, T, P))
        assert _near(kf[0], base(T), tol)
        assert kr[0] == 0.0
        assert _near(kf[1], -base(T), tol)          # negative A-factor
        assert _near(kf[2], 0.9 * base(T), tol)     # three lines at one pressure
    T, P = 3200.0, 1.0e6
    kf, kr = _k_by_reaction(o, T, _uniform_conc(o, T, P))
    for rx, sign in ((3, 1.0), (4, -1.0)):          # reverse from Keq
        assert _near(kf[rx], sign * base(T), tol)
        assert _near(kr[rx], sign * base(T) / _keq_h2(T), 1e-12)
    # three pressures with two lines each: 6.2, 1.3 and 10.01 at 1, 2 and 4 atm
    T = 3600.0
    cases = [(1.0e5, 6.2), (P_ATM, 6.2), (1.5 * P_ATM, 6.2 * 1.5 ** (np.log(1.3 / 6.2) / np.log(2.0))),
             (2.0 * P_ATM, 1.3), (3.5 * P_ATM, 1.3 * 1.75 ** (np.log(10.01 / 1.3) / np.log(2.0))),
             (4.0 * P_ATM, 10.01), (5.0 * P_ATM, 10.01)]
    for P, expect in cases:
        kf, _ = _k_by_reaction(o, T, _uniform_conc(o, T, P))
        assert _near(kf[5], expect, tol), (P, kf[5], expect)

    # TPGME-style table over 0.01-100 atm (plog_gtest.cpp:619-671), and the same
    # table with its A-factors split over several lines
    pts = np.array([0.01, 0.1, 1.0, 10.0, 100.0])
    lines = [(3.502e5, 1.441, -3244.0), (8.854e5, 1.327, -2975.0), (1.650e7, 0.973, -2010.0),
             (5.374e9, 0.287, 280.0), (9.494e18, -2.199, 9769.0)]

    def k7(P, T):
        pa = P / P_ATM
        kk = [A * T ** n * np.exp(-(E * E2T) / T) for A, n, E in lines]
        if pa < pts[0]:
            return kk[0] / 1000.0
        if pa >= pts[-1]:
            return kk[-1] / 1000.0
        i = int(np.nonzero((pts[:-1] <= pa) & (pa < pts[1:]))[0][0])
        lk = np.log(kk[i]) + np.log(kk[i + 1] / kk[i]) * np.log(pa / pts[i]) / np.log(pts[i + 1] / pts[i])
        return np.exp(lk) / 1000.0

    P, T = 1.0e3, 500.0
    while P < 2.0e7:
        kf, _ = _k_by_reaction(o, T, _uniform_conc(o, T, P))
        ref = k7(P, T)
        assert _near(kf[6], ref, 1e-12), (P, T)
        assert _near(kf[7], ref, 1e-12), (P, T)
        P *= 2.0
        T += 200.0


def test_ford_propane_closed_forms(oracle_mod):
    o = oracle_mod.Oracle(*mech_paths("ref_ford_propane"))
    t = o.tables
    idx = {s: j for j, s in enumerate(t.species)}
    fuel, oxid = 0.9028, 0.6855
    order = [fuel + oxid, 1.5, 1.0]
    x = np.array([0.233, 0.105, 0.206, 0.228, 0.121, 0.107])
    P = 10.0 * P_ATM

    def keq(T):
        lt = np.log(T)
        s = {"co": 3.5 * lt + 3.83145325, "o2": 3.5 * lt + 4.73253404,
             "co2": 4.46541338 * lt + 2.70309557e-01}
        h = {"co": 3.5 - 1.43372329e+04 / T, "o2": 3.5 - 1.04352500e+03 / T,
             "co2": 4.46541338 - 4.86597535e+04 / T}
        ds = -s["co"] - 0.5 * s["o2"] + s["co2"]
        dh = -h["co"] - 0.5 * h["o2"] + h["co2"]
        return (P_ATM / (RU * T)) ** -0.5 * np.exp(ds - dh)

    for T in (600.0, 800.0, 1000.0, 1200.0, 2000.0):
        C = x * P / (RU * T)
        k = np.array([2.0e12 * np.exp(-E2T * 3.3e3 / T), 4.5e10 * np.exp(-E2T * 1.2e3 / T), 0.0])
        k[0] *= 1000.0 ** (1.0 - order[0])
        k[1] *= 1000.0 ** (1.0 - order[1])
        k[2] = k[1] / keq(T)
        K = o.rate_constants(T, C)
        for s in range(3):
            assert _near(K[s], k[s], 1e-13), (T, s)
        rop = np.array([k[0] * C[idx["c3h8"]] ** fuel * C[idx["o2"]] ** oxid,
                        k[1] * C[idx["co"]] * C[idx["o2"]] ** 0.5, k[2] * C[idx["co2"]]])
        r, cr, de, net = o.rates(T, C, 0.0)
        for s in range(3):
            assert _near(r[s], rop[s], 1e-13), (T, s)
        cre = np.zeros(6)
        des = np.zeros(6)
        des[idx["c3h8"]] += rop[0]
        des[idx["o2"]] += 3.5 * rop[0] + 0.5 * rop[1]
        cre[idx["co"]] += 3.0 * rop[0] + rop[2]
        cre[idx["h2o"]] += 4.0 * rop[0]
        des[idx["co"]] += rop[1]
        cre[idx["co2"]] += rop[1]
        des[idx["co2"]] += rop[2]
        cre[idx["o2"]] += 0.5 * rop[2]
        for j in range(6):
            assert _near(cr[j], cre[j], 1e-13) and _near(de[j], des[j], 1e-13), (T, j)


def _frozen_density_rhs(o, y, rho0):
    t = o.tables
    Y, T = y[:-1], y[-1]
    C = rho0 * Y / t.mw
    _, _, _, net = o.rates(T, C, 0.0)
    cp_R, h_RT, _ = o.thermo(T)
    cp = np.sum(cp_R * RU / t.mw * Y)
    yd = np.empty_like(y)
    yd[:-1] = net * t.mw / rho0
    yd[-1] = -np.sum(h_RT * net / rho0) * RU * T / cp
    return yd


def _fd(f, y, j, h):
    cols = []
    for k in (-2, -1, 1, 2):
        yy = y.copy()
        yy[j] += k * h
        cols.append(f(yy))
    return (cols[0] - 8 * cols[1] + 8 * cols[2] - cols[3]) / (12 * h)


@pytest.mark.parametrize("name", ["ref_non_integer", "ref_ford_propane", "ref_rord_propane"])
@pytest.mark.parametrize("seed", [1, 2])
def test_nonint_jacobian_vs_finite_differences(oracle_mod, name, seed):
    o = oracle_mod.Oracle(*mech_paths(name))
    rng = np.random.default_rng(seed)
    n = o.nsp
    Y = rng.dirichlet(np.ones(n))
    T = rng.uniform(1200.0, 2000.0)
    P = rng.uniform(1.0, 10.0) * P_ATM
    y = np.append(Y, T)
    opts = oracle_mod.default_opts(step_limiter=0.0)
    Ja = o.jacobian(P, y, opts=opts)
    rho0 = P / (RU * T) / np.sum(Y / o.tables.mw)
    Jf = np.zeros((n + 1, n))
    for j in range(n):
        Jf[:, j] = _fd(lambda yy: _frozen_density_rhs(o, yy, rho0), y, j, 1e-5 * Y[j])
    scale = np.max(np.abs(Jf[:n]), axis=1, keepdims=True)
    err = np.abs(Ja[:n, :n] - Jf[:n])
    assert np.all(err <= 1e-6 * scale + 1e-300), np.max(err / (scale + 1e-300))
    # temperature row built from the species block
    tscale = np.max(np.abs(Jf[n]))
    assert np.all(np.abs(Ja[n, :n] - Jf[n]) <= 1e-6 * tscale)
