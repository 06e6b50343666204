"""The plugin's step callback on the GPU (zerork_reactor_set_callback_fn; the
manager sets it for a one-reactor solve with load_balance = 0,
zerork_reactor_manager.cpp:1082-1084, and solver_cvode.cpp:247-257 calls it
after every successful CVode call).  The integrator records each call on the
device and the plugin calls the callback over the records after the solve.
Checked on the one-wavefront, team and sparse paths:

* the callback changes nothing: the end state equals a solve without it bit
  for bit, and it is called once per CVode call (the COST output), with the
  reactor id the caller set;
* its times are the step ends (increasing, the last at or past dt), the
  differences add up to them, and the step sequence is the oracle's
  (zrko_integrate_one_cb) to within the usual FP64 divergence;
* its derivative is the interpolant's, zn[1] / h: against the oracle's
  right-hand side at the reported state within the corrector's
  Jacobian-propagated error bound (|J| + 1/h) w;
* a non-zero return ends the integration as a success with that record's
  state (bitwise), that record's call count as COST, and at constant volume
  the end pressure of that state, which the oracle's stop at the same call
  gives too.
"""
import numpy as np
import pytest

from conftest import mech_paths

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-6, 1.0e-12


def _plugin(capi, name, **opts):
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech_paths(name)) == 0
    base = dict(rel_tol=RTOL, abs_tol=ATOL, dense=1, iterative=0, analytic=1,
                constant_volume=0, fp32_preconditioner=0)
    base.update(opts)
    for k, v in base.items():
        rc = p.set_int(k, v) if isinstance(v, int) else p.set_double(k, v)
        assert rc == 0, k
    assert p.load_mechanism() == 0
    return p


def _solve(capi, name, st, k, dt, stop_at=None, callback=True, **opts):
    """One reactor through zerork_reactor_solve; returns (status, T, P, Y,
    cost, records, ids seen)."""
    p = _plugin(capi, name, load_balance=0 if callback else 1, **opts)
    nsp = st.Y.shape[1]
    recs, ids = [], []

    def cb(rid, nsteps, t, dtk, y, ydot, data):
        ids.append(rid)
        recs.append((nsteps, t, dtk, np.ctypeslib.as_array(y, (nsp + 1,)).copy(),
                     np.ctypeslib.as_array(ydot, (nsp + 1,)).copy()))
        return 1 if stop_at is not None and len(recs) - 1 == stop_at else 0

    assert p.set_callback(cb) == 0
    assert p.set_reactor_ids(np.array([7], dtype=np.int32)) == 0
    T, P, Y = st.T[k:k + 1].copy(), st.P[k:k + 1].copy(), st.Y[k:k + 1].copy()
    cost = np.zeros(1)
    assert p.set_aux("COST", cost) == 0
    rc = p.solve(0, 0.0, dt, T, P, Y)
    p.free()
    return rc, T[0], P[0], Y[0], cost[0], recs, ids


CASES = [("synth_10sp", "10sp", 1e-4, 2), ("synth_160sp", "160sp", None, 1),
         ("synth_2000sp", "2000sp", 1e-5, 0)]


@pytest.mark.parametrize("name,wl,dt,k", CASES)
def test_callback_matches_oracle_steps(torch_gpu, oracle_mod, name, wl, dt, k):
    from zerork_hip import capi, mechgen
    st = mechgen.workload_states(wl, mechgen.workload_mechanism(wl), 4, seed=23)
    dt = dt or mechgen.WORKLOAD_DT[wl]
    nsp = st.Y.shape[1]
    print("case callback", name, "reactor", k, "dt", dt, flush=True)
    rc, T, P, Y, cost, recs, ids = _solve(capi, name, st, k, dt)
    assert rc == 0
    rc0, T0, P0, Y0, cost0, recs0, _ = _solve(capi, name, st, k, dt, callback=False)
    assert rc0 == 0 and not recs0
    # the callback changes nothing and sees every CVode call
    assert T == T0
    np.testing.assert_array_equal(Y, Y0)
    assert cost == cost0 == len(recs) > 3
    assert set(ids) == {7}
    calls = np.array([r[0] for r in recs])
    np.testing.assert_array_equal(calls, np.arange(1, len(recs) + 1))
    t = np.array([r[1] for r in recs])
    assert np.all(np.diff(t) > 0) and t[-1] >= dt and (len(t) < 2 or t[-2] < dt)
    np.testing.assert_allclose(np.cumsum([r[2] for r in recs]), t, rtol=1e-12)
    if wl == "2000sp":
        return  # the oracle's dense LU takes minutes per reactor at this size
    o = oracle_mod.Oracle(*mech_paths(name))
    opts = oracle_mod.default_opts(rtol=RTOL, atol=ATOL, constant_volume=0)
    _, _, _, so, ro = o.integrate_steps(st.T[k], st.P[k], st.Y[k], dt, opts)
    to = np.array([r[1] for r in ro])
    m = min(len(t), len(to))
    same = int(np.argmax(np.abs(t[:m] - to[:m]) > 1e-6 * to[:m])) if np.any(
        np.abs(t[:m] - to[:m]) > 1e-6 * to[:m]) else m
    print(f"{wl}: records gpu {len(t)} oracle {len(to)}; step times agree (1e-6) over the "
          f"first {same}", flush=True)
    assert abs(len(t) - len(to)) <= max(2, 0.05 * len(to))
    assert same >= min(10, m)
    # the derivative against the oracle's right-hand side at the reported state
    worst = 0.0
    for _, _, h, y, yd in recs:
        f, flag = o.rhs(st.P[k], y, opts)
        assert flag == 0
        J = o.jacobian(st.P[k], y, f, opts)
        w = RTOL * np.abs(y) + ATOL
        worst = max(worst, float(np.max(np.abs(yd - f) / (np.abs(J) @ w + w / h))))
    print(f"{wl}: max |ydot - f| / ((|J| + 1/h) w) = {worst:.3e}", flush=True)
    assert worst < 4.0 * np.sqrt(nsp + 1)


@pytest.mark.parametrize("name,wl,cvol", [("synth_10sp", "10sp", 1), ("synth_160sp", "160sp", 0),
                                          ("synth_2000sp", "2000sp", 0)])
def test_callback_stop_ends_the_solve(torch_gpu, oracle_mod, name, wl, cvol):
    from zerork_hip import capi, mechgen
    st = mechgen.workload_states(wl, mechgen.workload_mechanism(wl), 4, seed=23)
    dt = 1e-5 if wl == "2000sp" else mechgen.WORKLOAD_DT[wl]
    k, stop = 1, 5
    print("case callback stop", name, "constant volume", cvol, flush=True)
    rc, _, _, _, _, full, _ = _solve(capi, name, st, k, dt, constant_volume=cvol)
    assert rc == 0 and len(full) > stop + 1
    rc, T, P, Y, cost, recs, _ = _solve(capi, name, st, k, dt, stop_at=stop, constant_volume=cvol)
    assert rc == 0 and len(recs) == stop + 1 and cost == stop + 1
    for a, b in zip(recs, full):
        assert a[:3] == b[:3]
        np.testing.assert_array_equal(a[3], b[3])
    y = recs[stop][3]
    np.testing.assert_array_equal(Y, y[:-1])
    assert T == y[-1]
    if not cvol:
        assert P == st.P[k]
    else:
        o = oracle_mod.Oracle(*mech_paths(name))
        opts = oracle_mod.default_opts(rtol=RTOL, atol=ATOL, constant_volume=1)
        To, Po, Yo, so, ro = o.integrate_steps(st.T[k], st.P[k], st.Y[k], dt, opts,
                                               stop_at=stop)
        assert so["nsteps"] == stop + 1
        print(f"end pressure at the stop: gpu {P:.12e} oracle {Po:.12e}", flush=True)
        assert abs(P - Po) <= 1e-9 * Po
        assert abs(T - To) <= 1e-6 * To
