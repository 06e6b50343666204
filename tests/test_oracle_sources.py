"""The oracle's restatement of the plugin's source terms (DPDT, E_SRC, Y_SRC:
reactor_constant_pressure_cpu.cpp:15-133, reactor_constant_volume_cpu.cpp:
15-130), pinned by what the equations fix:

* zero-valued source arrays give bit-identical results to no sources;
* constant pressure, energy source only: the mass enthalpy rises as e_src t
  (dh/dt = e_src + v dP/dt for the reactor's equations, for any chemistry),
  so h(T_end, Y_end) - h0 = e_src dt, to the accuracy with which the RHS's
  temperature (solved from the energy, tolerance 1e-6 K) follows the state's;
* mass sources: sum(Ydot) = sum(y_src) because sum(MW_j wdot_j) = 0, so
  sum(Y_end) - 1 = sum(y_src) dt (a linear invariant the BDF keeps up to its
  Newton iterations with the approximate Jacobian), and they leave the mass
  enthalpy unchanged (dh/dt = 0);
* pressure rate: the end pressure is P0 + dpdt dt, and the mass enthalpy gains
  the integral of v dpdt (checked by the trapezoidal rule over the step);
* the energy-to-temperature solve: at t = 0 the source-term RHS equals the
  plain RHS plus e_src / cp in the temperature row (the solve returns the
  state's own temperature to its 1e-6 K tolerance).
* the constant-volume pressure rate (reactor_constant_volume_cpu.cpp:94-98):
  at time t the reactor of fixed volume v0 evaluates its rates at the
  density of the pressure Ru T/(v0 W) + dpdt t, so its RHS equals the plain
  constant-volume RHS of a reactor whose volume gives that pressure, plus
  dpdt v_t / cv in the temperature row (v_t the volume of that density); at
  t = 0 this is the plain RHS of the reactor itself plus dpdt v0 / cv.  A
  density taken at v0, a dropped dpdt t or a missing energy term breaks it.
"""
import numpy as np
import pytest

from conftest import mech_paths

RU = 8.314462618e3
RTOL, ATOL = 1.0e-6, 1.0e-12


def _states(n, seed):
    from zerork_hip import mechgen
    st = mechgen.workload_states("10sp", mechgen.workload_mechanism("10sp"), n, seed=seed)
    return st.T, st.P, st.Y


def _h_mass(o, T, Y, cvol=False):
    mw = o.tables.mw
    out = np.empty(len(T))
    for k in range(len(T)):
        _, h, _ = o.thermo(T[k])
        h = h - (1.0 if cvol else 0.0)
        out[k] = np.sum(h * (RU / mw) * T[k] * Y[k])
    return out


@pytest.fixture(scope="module")
def o10(oracle_mod):
    return oracle_mod.Oracle(*mech_paths("synth_10sp"))


def test_zero_sources_are_bit_identical(oracle_mod, o10):
    T, P, Y = _states(16, 3)
    oo = oracle_mod.default_opts(rtol=RTOL, atol=ATOL)
    a = o10.integrate(T, P, Y, 1e-4, oo)
    b = o10.integrate(T, P, Y, 1e-4, oo, dpdt=np.zeros(16), esrc=np.zeros(16))
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2]["nsteps"], b[2]["nsteps"])


def test_energy_source_raises_enthalpy_by_esrc_dt(oracle_mod, o10):
    T, P, Y = _states(24, 4)
    dt = 1e-4
    esrc = np.linspace(-2e6, 5e6, 24)  # J/kg/s
    oo = oracle_mod.default_opts(rtol=1e-9, atol=1e-15)
    Te, Ye, st = o10.integrate(T, P, Y, dt, oo, esrc=esrc)
    assert np.all(st["nsteps"] > 0)
    dh = _h_mass(o10, Te, Ye) - _h_mass(o10, T, Y)
    np.testing.assert_allclose(dh, esrc * dt, rtol=2e-3, atol=1e-2)
    np.testing.assert_array_equal(st["P"], P)


def test_mass_sources_shift_sum_and_keep_enthalpy(oracle_mod, o10):
    T, P, Y = _states(24, 5)
    dt = 1e-4
    rng = np.random.default_rng(6)
    # relaxation towards another composition plus a net inflow (1/s)
    ys = 50.0 * (rng.dirichlet(np.ones(Y.shape[1]), Y.shape[0]) - Y) + 2.0 * Y
    oo = oracle_mod.default_opts(rtol=1e-9, atol=1e-15)
    Te, Ye, st = o10.integrate(T, P, Y, dt, oo, ysrc=ys)
    assert np.all(st["nsteps"] > 0)
    np.testing.assert_allclose(Ye.sum(axis=1) - 1.0, ys.sum(axis=1) * dt, rtol=0, atol=1e-10)
    h0, h1 = _h_mass(o10, T, Y), _h_mass(o10, Te, Ye)
    assert np.all(np.abs(h1 - h0) <= 1e-6 * np.abs(h0) + 1.0)


def test_pressure_rate(oracle_mod, o10):
    T, P, Y = _states(24, 7)
    dt = 1e-4
    dpdt = np.linspace(-1e8, 3e8, 24)  # Pa/s
    oo = oracle_mod.default_opts(rtol=1e-9, atol=1e-15)
    Te, Ye, st = o10.integrate(T, P, Y, dt, oo, dpdt=dpdt)
    assert np.all(st["nsteps"] > 0)
    np.testing.assert_allclose(st["P"], P + dpdt * dt, rtol=1e-15)
    # dh/dt = v dpdt: 50 chained sub-steps (each restarts at t = 0 from the
    # previous end state and pressure, which continues P(t) = P0 + dpdt t),
    # with the trapezoidal rule for v over each
    mw = o10.tables.mw
    v = lambda TT, PP, YY: RU * TT * np.sum(YY / mw, axis=1) / PP
    Tk, Pk, Yk = T.copy(), P.copy(), Y.copy()
    expect = np.zeros(len(T))
    for _ in range(50):
        T1, Y1, s1 = o10.integrate(Tk, Pk, Yk, dt / 50, oo, dpdt=dpdt)
        expect += dpdt * (dt / 50) * 0.5 * (v(Tk, Pk, Yk) + v(T1, s1["P"], Y1))
        Tk, Pk, Yk = T1, s1["P"], Y1
    np.testing.assert_allclose(Pk, P + dpdt * dt, rtol=1e-12)
    dh = _h_mass(o10, Tk, Yk) - _h_mass(o10, T, Y)
    np.testing.assert_allclose(dh, expect, rtol=2e-3, atol=1e-2)


@pytest.mark.parametrize("cvol", [0, 1])
def test_energy_source_rhs_at_t0(oracle_mod, o10, cvol):
    T, P, Y = _states(8, 8)
    oo = oracle_mod.default_opts(constant_volume=cvol)
    for k in range(8):
        y = np.append(Y[k], T[k])
        f0, _ = o10.rhs(P[k], y, oo)
        f1, _ = o10.rhs_src(P[k], y, 0.0, y, esrc=3e6, opts=oo)
        np.testing.assert_allclose(f1[:-1], f0[:-1], rtol=1e-6, atol=1e-9 * np.max(np.abs(f0[:-1])))
        cp_R, _, _ = o10.thermo(T[k])
        cx = np.sum((cp_R - (1.0 if cvol else 0.0)) * RU / o10.tables.mw * Y[k])
        np.testing.assert_allclose(f1[-1] - f0[-1], 3e6 / cx, rtol=1e-6)


@pytest.mark.parametrize("t", [0.0, 1e-4, 3e-3])
def test_constant_volume_pressure_rate_rhs(oracle_mod, o10, t):
    T, P, Y = _states(8, 9)
    oo = oracle_mod.default_opts(constant_volume=1)
    mw = o10.tables.mw
    rng = np.random.default_rng(10)
    for k in range(8):
        y0 = np.append(Y[k], T[k])
        # the state at time t differs from the initial one (T and Y moved)
        Yt = Y[k] * (1.0 + 0.2 * rng.standard_normal(len(mw)))
        Yt = np.abs(Yt) / np.sum(np.abs(Yt))
        Tt = T[k] * 1.07
        y = np.append(Yt, Tt)
        dpdt = 3.0e8 if k % 2 == 0 else -1.5e8
        f1, fl = o10.rhs_src(P[k], y0, t, y, dpdt=dpdt, opts=oo)
        assert fl == 0
        v0 = RU * T[k] * np.sum(Y[k] / mw) / P[k]        # the reactor's fixed volume
        Pt = RU * Tt * np.sum(Yt / mw) / v0 + dpdt * t    # its pressure at t
        vt = RU * Tt * np.sum(Yt / mw) / Pt              # the volume of that density
        f0, fl = o10.rhs(Pt, y, oo)                      # plain RHS at volume vt
        assert fl == 0
        cp_R, _, _ = o10.thermo(Tt)
        cv = np.sum((cp_R - 1.0) * RU / mw * Yt)
        np.testing.assert_allclose(f1[:-1], f0[:-1], rtol=1e-10,
                                   atol=1e-12 * np.max(np.abs(f0[:-1])))
        np.testing.assert_allclose(f1[-1], f0[-1] + dpdt * vt / cv, rtol=1e-10,
                                   atol=1e-12 * abs(dpdt * vt / cv))
