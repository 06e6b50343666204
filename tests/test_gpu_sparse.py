"""GPU parity of the sparse path (one reactor per workgroup, static sparse
FP32 LU of the Newton matrix I - gamma J: the reference's sparse direct
solve, reactor_nvector_serial.cpp:414-507 with superlu_manager.cpp:13-192),
through zrk_integrate_device.

* Forced onto the 10-, 53- and 160-species mechanisms, constant pressure and
  constant volume: end states against the oracle (dense FP64 LU) in
  tolerance units, within the bounds the FP32 dense path meets (the oracle's
  own one-ulp sensitivity plus an allowance for the FP32 Newton matrix,
  twice the dense FP32 path's at the 90th percentile: the sparse LU
  eliminates in another order, so its round-off differs), median step
  count, sum(Y), the end pressure at constant volume, and accuracy against
  the LSODA referee on a subset (no worse than the oracle's own distance
  from it, plus the allowance the FP32 dense path has).
* 2,000 species / 8,000 reactions, the large workload at its full size
  (4,096 reactors over dt = 1e-4 s): four reactors whose oracle end states
  are committed (tests/golden/sparse_2000sp_fixture.npz, written by
  make_sparse_fixture.py with the oracle's own one-ulp sensitivity beside
  them; the oracle needs about a minute per reactor here) sit at four
  places in the batch and meet the FP32 bar above against them; over the
  whole batch every reactor succeeds, sum(Y) = 1, the adiabatic
  constant-pressure invariant holds (the mixture's mass enthalpy keeps its
  initial value to the integration tolerance), and the fixture reactors
  integrated alone give bitwise the batch's results.
* Options the path does not take are refused with the documented codes.
* The drop-in plugin takes the sparse path for the 2,000-species mechanism by
  itself, with the device entry point's results bit for bit.
"""
import time

import numpy as np
import pytest

from conftest import mech_paths
from test_gpu_integrate import _gpu_integrate, _referee, _units

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-6, 1.0e-12
RU = 8.314462618e3


@pytest.mark.parametrize("wl,n,cvol", [("10sp", 256, 0), ("10sp", 256, 1), ("53sp", 192, 0),
                                       ("53sp", 128, 1), ("160sp", 48, 0)])
def test_sparse_path_matches_oracle(torch_gpu, oracle_mod, wl, n, cvol):
    from zerork_hip import capi, mechgen
    name = "synth_" + wl
    m = capi.Mech(*mech_paths(name))
    m.set_sparse(True)
    o = oracle_mod.Oracle(*mech_paths(name))
    st = mechgen.workload_states(wl, mechgen.workload_mechanism(wl), n, seed=61)
    dt = mechgen.WORKLOAD_DT[wl]
    print("case sparse", wl, n, "constant volume", cvol, flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, constant_volume=cvol,
                                fp32_preconditioner=1)
    Tg, Yg, steps, stats, Pg = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, dt, opts, True)
    oo = oracle_mod.default_opts(rtol=RTOL, atol=ATOL, constant_volume=cvol)
    To, Yo, so = o.integrate(st.T, st.P, st.Y, dt, oo)
    Tp, Yp, sp = o.integrate(np.nextafter(st.T, np.inf), st.P, st.Y, dt, oo)
    assert np.all(steps > 0) and np.all(stats[:, 7] == 0)
    gpu = _units(Yg, Tg, Yo, To, RTOL, ATOL)
    ulp = _units(Yp, Tp, Yo, To, RTOL, ATOL)
    q = [50, 90, 100]
    gq, uq = np.percentile(gpu, q), np.percentile(ulp, q)
    print(f"{wl}: sparse-oracle tolerance units p50/p90/max {gq.round(2)}; oracle 1-ulp "
          f"{uq.round(2)}; median steps gpu {np.median(steps)} oracle "
          f"{np.median(so['nsteps'])}", flush=True)
    assert gq[0] <= 2.0 * uq[0] + 1.0 and gq[1] <= 2.0 * uq[1] + 20.0
    assert gq[2] <= 4.0 * uq[2] + 50.0
    dmed = abs(np.median(sp["nsteps"]) / np.median(so["nsteps"]) - 1.0)
    assert abs(np.median(steps) / np.median(so["nsteps"]) - 1.0) < 0.05 + 2.0 * dmed
    assert np.all(np.abs(Yg.sum(axis=1) - 1.0) <= 1e-8)
    if cvol:
        pu = np.abs(Pg - so["P"]) / (RTOL * so["P"])
        assert np.percentile(pu, 90) <= 2.0 * np.percentile(ulp, 90) + 10.0
    else:
        np.testing.assert_array_equal(Pg, st.P)
    sub = np.linspace(0, n - 1, 5).astype(int)
    eg, eo = [], []
    for i in sub:
        yr, tr = _referee(o, st.T[i], st.P[i], st.Y[i], dt, oo)
        eg.append(_units(Yg[i:i + 1], Tg[i:i + 1], yr[None], np.array([tr]), RTOL, ATOL)[0])
        eo.append(_units(Yo[i:i + 1], To[i:i + 1], yr[None], np.array([tr]), RTOL, ATOL)[0])
    print(f"{wl}: error vs referee (tol units) sparse {np.round(eg, 1)} oracle {np.round(eo, 1)}",
          flush=True)
    assert max(eg) <= 2.0 * max(eo) + 50.0


def _mass_enthalpy(o, T, Y):
    mw = o.tables.mw
    h = np.array([o.thermo(t)[1] for t in T])  # h/RT per species
    return np.sum(h * RU * T[:, None] / mw * Y, axis=1)


def test_sparse_path_2000_species_workload(torch_gpu, oracle_mod):
    """The 2,000-species / 8,000-reaction workload at its full size (4,096
    constant-pressure reactors over dt = 1e-4 s) on the sparse path: four
    reactors against committed FP64 oracle end states (within four times
    the oracle's own one-ulp sensitivity plus 50 units, step counts within
    10 %), the same reactors alone bitwise equal to the batch, sum(Y) = 1 and
    the adiabatic enthalpy invariant.  The step cap is 1,500 CVode steps with
    no retry: a reactor that needs more ends with CV_TOO_MUCH_WORK and keeps
    its input state, as the plugin's solver does (solver_cvode.cpp:228-244);
    at most 1 % of the batch may: about 25 of these 4,096 reactors (which
    ones moves with the round-off of the workgroup's reductions) need more
    than 1,500 steps or stop at a Newton convergence failure, the others at
    most about 950, and so does the FP64 oracle on the one checked (reactor
    397: 1,500 of 1,500 steps, 6 convergence failures)."""
    from zerork_hip import capi, mechgen
    from conftest import GOLDEN
    import os
    name = "synth_2000sp"
    m = capi.Mech(*mech_paths(name))
    nnz, lev, flops = m.sparse_plan_info()
    print(f"case sparse 2000sp workload: LU entries {nnz}, levels {lev}, multiply-adds {flops}",
          flush=True)
    with np.load(os.path.join(GOLDEN, "sparse_2000sp_fixture.npz"), allow_pickle=False) as z:
        fx = {k: z[k] for k in z.files}
    n = mechgen.WORKLOAD_NR["2000sp"]
    dt = float(fx["dt"])
    st = mechgen.workload_states("2000sp", mechgen.workload_mechanism("2000sp"), n, seed=2)
    at = np.array([0, n // 3, (2 * n) // 3, n - 1])
    st.T[at], st.P[at], st.Y[at] = fx["T0"], fx["P0"], fx["Y0"]
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, constant_volume=0,
                                fp32_preconditioner=1, max_steps=1500, cvode_num_retries=1)
    t0 = time.perf_counter()
    Tg, Yg, steps, stats = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, dt, opts)
    el = time.perf_counter() - t0
    ok = steps > 0
    print(f"2000sp: {n} reactors in {el:.1f} s; steps mean {steps[ok].mean():.1f} max "
          f"{steps[ok].max()}; setups {stats[ok, 3].mean():.1f}; at the step cap "
          f"{np.flatnonzero(~ok).tolist()}", flush=True)
    assert np.sum(~ok) <= 0.01 * n
    # CV_TOO_MUCH_WORK at the cap, or CV_CONV_FAILURE (the FP64 oracle also
    # meets Newton convergence failures on these reactors)
    assert np.all(stats[ok, 7] == 0) and np.all(np.isin(stats[~ok, 7], (-1, -4)))
    assert np.sum(stats[~ok, 7] == -4) <= 2
    np.testing.assert_array_equal(Tg[~ok], st.T[~ok])
    np.testing.assert_array_equal(Yg[~ok], st.Y[~ok])
    assert np.all(ok[at])
    # the fixture reactors against the oracle, in tolerance units
    gpu = _units(Yg[at], Tg[at], fx["Y"], fx["T"], RTOL, ATOL)
    ulp = _units(fx["Y_ulp"], fx["T_ulp"], fx["Y"], fx["T"], RTOL, ATOL)
    print(f"2000sp fixture: GPU-oracle tolerance units {gpu.round(2)}; oracle 1-ulp "
          f"{ulp.round(2)}; steps gpu {steps[at]} oracle {fx['nsteps']}", flush=True)
    assert np.all(gpu <= 4.0 * np.maximum(ulp, 1.0) + 50.0)
    assert abs(steps[at].mean() / fx["nsteps"].mean() - 1.0) < 0.1
    # the same reactors alone: bitwise the batch's results
    Ta, Ya, sa, _ = _gpu_integrate(torch_gpu, m, fx["T0"], fx["P0"], fx["Y0"], dt, opts)
    np.testing.assert_array_equal(Ta, Tg[at])
    np.testing.assert_array_equal(Ya, Yg[at])
    np.testing.assert_array_equal(sa, steps[at])
    # whole batch: sum(Y) and the adiabatic constant-pressure invariant
    assert np.all(np.abs(Yg[ok].sum(axis=1) - 1.0) <= 1e-8)
    o = oracle_mod.Oracle(*mech_paths(name))
    sub = np.linspace(0, n - 1, 256).astype(int)
    sub = sub[ok[sub]]
    h0, h1 = _mass_enthalpy(o, st.T[sub], st.Y[sub]), _mass_enthalpy(o, Tg[sub], Yg[sub])
    cp = np.array([np.sum(o.thermo(t)[0] * RU / o.tables.mw * y)
                   for t, y in zip(Tg[sub], Yg[sub])])
    drift = np.abs(h1 - h0) / cp
    print(f"2000sp: enthalpy drift / cp max {drift.max():.3e} K (256 reactors)", flush=True)
    assert np.all(drift <= 50.0 * (RTOL * Tg[sub] + ATOL))


def test_sparse_path_refuses_unsupported(torch_gpu):
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_10sp"))
    m.set_sparse(True)
    st = mechgen.workload_states("10sp", mechgen.workload_mechanism("10sp"), 4, seed=63)
    dev = torch.device("cuda:0")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    for kw, code in ((dict(fp32_preconditioner=0), -3),
                     (dict(fp32_preconditioner=1, delta_temperature_ignition=400.0), -4)):
        with pytest.raises(RuntimeError, match=rf"\({code}\)"):
            m.integrate(4, 1e-4, t(st.T), t(st.P), t(st.Y), 10, capi.default_options(**kw))


def test_plugin_runs_large_mechanism_on_sparse_path(torch_gpu):
    """The drop-in plugin (zerork_reactor_solve) takes the sparse path for a
    2,000-species mechanism by itself and gives the device entry point's
    results bit for bit."""
    from zerork_hip import capi, mechgen
    name = "synth_2000sp"
    st = mechgen.workload_states("2000sp", mechgen.workload_mechanism("2000sp"), 4, seed=64)
    n, nsp = st.Y.shape
    dt = 1.0e-7
    print("case plugin sparse 2000sp", flush=True)
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech_paths(name)) == 0
    for k, v in (("constant_volume", 0), ("rel_tol", RTOL), ("abs_tol", ATOL)):
        assert (p.set_int(k, v) if isinstance(v, int) else p.set_double(k, v)) == 0, k
    assert p.load_mechanism() == 0
    T, P, Y = st.T.copy(), st.P.copy(), st.Y.copy()
    cost = np.zeros(n)
    assert p.set_aux("COST", cost) == 0
    assert p.solve(0, 0.0, dt, T, P, Y) == 0
    m = capi.Mech(*mech_paths(name))
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, constant_volume=0,
                                fp32_preconditioner=1)
    Td, Yd, sd, _ = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, dt, opts)
    assert np.all(sd > 0)
    np.testing.assert_array_equal(T, Td)
    np.testing.assert_array_equal(Y, Yd)
    np.testing.assert_array_equal(cost, sd)
