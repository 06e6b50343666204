"""The oracle's step callback (the plugin's zerork_reactor_set_callback_fn:
solver_cvode.cpp:228-272 calls it after every successful CVode call of a
one-reactor solve with the state, CVodeGetDky(t, 1) and the time since the
previous call; a non-zero return ends the integration there), pinned by:

* the callback observes and changes nothing: the end state of a run with a
  callback that returns 0 is bitwise the batch integrator's, and it is
  called once per CVode call (the cost output);
* the times are CVode's internal step ends: increasing, the last one at or
  past dt and the one before short of it; the differences add up to the last;
* the derivative is the BDF interpolant's at the step end, zn[1] / h, which the
  converged corrector makes f(t_n, y_n) up to the corrector's own error
  delta (|delta| of the order of CVODE's error weights w): compared with the
  right-hand side at the reported state against (|J| + 1/h) w, the
  Jacobian-propagated bound, and in the median relative to |f| (a missing
  1/h fails by the step's size, orders of magnitude);
* a non-zero return ends the integration with that record's state (bitwise)
  and its call count as the cost; on the last record (t past dt) the state is
  the interpolant at dt, the normal end state;
* with the ignition root the root returns are records too: their times are
  the ignition time that the batch integrator reports.
"""
import numpy as np
import pytest

from conftest import mech_paths


def _case(wl, n, seed):
    from zerork_hip import mechgen
    mg = mechgen.workload_mechanism(wl)
    st = mechgen.workload_states(wl, mg, n, seed=seed)
    return st, mechgen.WORKLOAD_DT[wl]


@pytest.mark.parametrize("wl,k,cvol", [("10sp", 0, 0), ("10sp", 3, 1), ("53sp", 1, 0)])
def test_callback_sees_every_step_and_changes_nothing(oracle_mod, wl, k, cvol):
    o = oracle_mod.Oracle(*mech_paths("synth_" + wl))
    st, dt = _case(wl, 4, 7)
    opts = oracle_mod.default_opts(rtol=1e-6, atol=1e-12, constant_volume=cvol)
    Tb, Yb, sb = o.integrate(st.T[k:k + 1], st.P[k:k + 1], st.Y[k:k + 1], dt, opts)
    T, P, Y, s, recs = o.integrate_steps(st.T[k], st.P[k], st.Y[k], dt, opts)
    assert T == Tb[0] and P == sb["P"][0]
    np.testing.assert_array_equal(Y, Yb[0])
    assert s["nsteps"] == sb["nsteps"][0] == len(recs) > 5
    calls = np.array([r[0] for r in recs])
    np.testing.assert_array_equal(calls, np.arange(1, len(recs) + 1))
    t = np.array([r[1] for r in recs])
    assert np.all(np.diff(t) > 0) and t[-1] >= dt > t[-2]
    dts = np.array([r[2] for r in recs])
    assert dts[0] == t[0]
    np.testing.assert_allclose(np.cumsum(dts), t, rtol=1e-12)
    assert all(r[3].shape == (o.nsp + 1,) and r[4].shape == (o.nsp + 1,) for r in recs)


@pytest.mark.parametrize("wl", ["10sp", "53sp"])
def test_callback_derivative_is_the_right_hand_side(oracle_mod, wl):
    o = oracle_mod.Oracle(*mech_paths("synth_" + wl))
    st, dt = _case(wl, 3, 7)
    rtol, atol = 1e-6, 1e-12
    opts = oracle_mod.default_opts(rtol=rtol, atol=atol, constant_volume=0)
    worst, rel = 0.0, []
    for k in range(3):
        _, _, _, _, recs = o.integrate_steps(st.T[k], st.P[k], st.Y[k], dt, opts)
        for _, _, h, y, yd in recs:
            f, flag = o.rhs(st.P[k], y, opts)
            assert flag == 0
            J = o.jacobian(st.P[k], y, f, opts)
            w = rtol * np.abs(y) + atol
            bound = np.abs(J) @ w + w / h
            worst = max(worst, float(np.max(np.abs(yd - f) / bound)))
            rel.append(np.median(np.abs(yd - f) / (np.abs(f) + 1e-3 * np.abs(f).max())))
    n = o.nsp + 1
    print(f"{wl}: max |ydot - f| / ((|J| + 1/h) w) = {worst:.3e} (bar {4 * np.sqrt(n):.1f}); "
          f"median relative {np.median(rel):.2e}")
    # per component against an rms-norm tolerance: up to sqrt(N) times it
    assert worst < 4.0 * np.sqrt(n)
    assert np.median(rel) < 1e-6


@pytest.mark.parametrize("stop", [0, 4, "last"])
def test_callback_stop_ends_the_integration(oracle_mod, stop):
    o = oracle_mod.Oracle(*mech_paths("synth_10sp"))
    st, dt = _case("10sp", 2, 7)
    opts = oracle_mod.default_opts(rtol=1e-6, atol=1e-12, constant_volume=0)
    Tf, Pf, Yf, sf, full = o.integrate_steps(st.T[1], st.P[1], st.Y[1], dt, opts)
    k = len(full) - 1 if stop == "last" else stop
    T, P, Y, s, recs = o.integrate_steps(st.T[1], st.P[1], st.Y[1], dt, opts, stop_at=k)
    assert len(recs) == k + 1 and s["nsteps"] == k + 1 and s["status"] == 0
    for a, b in zip(recs, full):  # the same calls up to the stop
        assert a[:3] == b[:3]
        np.testing.assert_array_equal(a[3], b[3])
    if stop == "last":
        assert T == Tf
        np.testing.assert_array_equal(Y, Yf)
    else:
        y = recs[k][3]
        np.testing.assert_array_equal(Y, y[:-1])
        assert T == y[-1] * opts.Tref
        assert P == st.P[1]  # constant pressure, no sources


def test_callback_sees_root_returns(oracle_mod):
    o = oracle_mod.Oracle(*mech_paths("synth_10sp"))
    st, dt = _case("10sp", 24, 11)
    opts = oracle_mod.default_opts(rtol=1e-6, atol=1e-12, delta_temperature_ignition=400.0)
    _, _, sb = o.integrate(st.T, st.P, st.Y, dt, opts)
    ign = sb["ignition_time"]
    k = int(np.argmax(ign > 0))
    assert ign[k] > 0
    _, _, _, s, recs = o.integrate_steps(st.T[k], st.P[k], st.Y[k], dt, opts)
    t = np.array([r[1] for r in recs])
    assert len(recs) == s["nsteps"] == sb["nsteps"][k]
    assert np.count_nonzero(t == ign[k]) == 1
    j = int(np.flatnonzero(t == ign[k])[0])
    # the root return's state is on the interpolant: T there is the
    # ignition temperature to within the root tolerance
    Tj = recs[j][3][-1] * opts.Tref
    assert abs(Tj - (st.T[k] + 400.0)) < 1e-3
