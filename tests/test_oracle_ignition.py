"""The oracle's ignition-time root finding (the plugin's
delta_temperature_ignition / IGNITION_TIME / stop_after_ignition:
reactor_nvector_serial.cpp:774-810 and solver_cvode.cpp:90-92, 228-270,
restating SUNDIALS CVODE 5.8.0's cvRcheck1-3 and cvRootfind), pinned by:

* CVODE locates roots on the interpolant of steps it takes anyway: the end
  states are bitwise those of the run without root finding, and each crossing
  costs exactly one more CVode call (the root return) in the cost output;
* convergence to an independent referee: the crossing time of T0 + delta_T
  found by scipy's LSODA event location at rtol 1e-12; the oracle's error
  must shrink with its tolerance (one-sided check at rtol 1e-10, where it is
  a few 1e-9 relative);
* stop_after_ignition returns the state at the crossing: T = T0 + delta_T to
  within the root tolerance, and the same ignition time;
* reactors that do not reach the ignition temperature report 0.
"""
import numpy as np
import pytest

from conftest import mech_paths

DT_IGN = 400.0


def _case(wl, n, seed):
    from zerork_hip import mechgen
    mg = mechgen.workload_mechanism(wl)
    st = mechgen.workload_states(wl, mg, n, seed=seed)
    return st, mechgen.WORKLOAD_DT[wl]


@pytest.mark.parametrize("wl,cvol", [("10sp", 0), ("10sp", 1), ("53sp", 0)])
def test_root_finding_leaves_the_steps_alone(oracle_mod, wl, cvol):
    o = oracle_mod.Oracle(*mech_paths("synth_" + wl))
    st, dt = _case(wl, 48 if wl == "10sp" else 24, 11)
    base = dict(rtol=1e-6, atol=1e-12, constant_volume=cvol)
    T0, Y0, s0 = o.integrate(st.T, st.P, st.Y, dt, oracle_mod.default_opts(**base))
    T1, Y1, s1 = o.integrate(st.T, st.P, st.Y, dt, oracle_mod.default_opts(
        delta_temperature_ignition=DT_IGN, **base))
    ign = s1["ignition_time"]
    assert np.any(ign > 0) and np.any(ign == 0)
    np.testing.assert_array_equal(T1, T0)
    np.testing.assert_array_equal(Y1, Y0)
    np.testing.assert_array_equal(s1["nst"], s0["nst"])
    # one root return per crossing (T rises through T0 + delta_T once)
    np.testing.assert_array_equal(s1["nsteps"] - s0["nsteps"], (ign > 0).astype(int))
    # reactors that ignite end above the ignition temperature, the others below
    assert np.all(T1[ign > 0] > st.T[ign > 0] + DT_IGN)
    assert np.all(T1[ign == 0] < st.T[ign == 0] + DT_IGN)
    assert np.all((ign[ign > 0] > 0) & (ign[ign > 0] <= dt))


def test_ignition_time_converges_to_referee(oracle_mod):
    from scipy.integrate import solve_ivp
    o = oracle_mod.Oracle(*mech_paths("synth_10sp"))
    st, dt = _case("10sp", 24, 11)
    errs = {}
    for rtol in (1e-6, 1e-10):
        _, _, s = o.integrate(st.T, st.P, st.Y, dt, oracle_mod.default_opts(
            rtol=rtol, atol=rtol * 1e-6, delta_temperature_ignition=DT_IGN))
        errs[rtol] = s["ignition_time"]
    ids = np.nonzero(errs[1e-10] > 0)[0][:8]
    assert len(ids) >= 4
    e6, e10 = [], []
    for i in ids:
        Tign = st.T[i] + DT_IGN
        ev = lambda t, y: Tign - y[-1]
        sol = solve_ivp(lambda t, y: o.rhs(st.P[i], y)[0], (0.0, dt), np.append(st.Y[i], st.T[i]),
                        method="LSODA", rtol=1e-12, atol=1e-24, events=ev)
        tr = sol.t_events[0][0]
        e6.append(abs(errs[1e-6][i] - tr) / tr)
        e10.append(abs(errs[1e-10][i] - tr) / tr)
    print("relative error vs referee: rtol 1e-6", np.max(e6), "rtol 1e-10", np.max(e10))
    assert max(e10) <= 1e-7
    assert max(e10) <= 0.1 * max(e6)


@pytest.mark.parametrize("cvol", [0, 1])
def test_stop_after_ignition_returns_the_crossing_state(oracle_mod, cvol):
    o = oracle_mod.Oracle(*mech_paths("synth_10sp"))
    st, dt = _case("10sp", 32, 12)
    base = dict(rtol=1e-6, atol=1e-12, constant_volume=cvol, delta_temperature_ignition=DT_IGN)
    _, _, s1 = o.integrate(st.T, st.P, st.Y, dt, oracle_mod.default_opts(**base))
    T2, Y2, s2 = o.integrate(st.T, st.P, st.Y, dt, oracle_mod.default_opts(
        stop_after_ignition=1, **base))
    ign = s1["ignition_time"]
    m = ign > 0
    assert m.any()
    np.testing.assert_array_equal(s2["ignition_time"], ign)
    assert np.all(np.abs(T2[m] - (st.T[m] + DT_IGN)) <= 1e-9 * (st.T[m] + DT_IGN))
    # stopped reactors took no steps past the crossing
    assert np.all(s2["nst"][m] <= s1["nst"][m])
    np.testing.assert_allclose(Y2.sum(axis=1), 1.0, rtol=0, atol=1e-9)
