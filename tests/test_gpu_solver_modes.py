"""GPU parity of the plugin's linear-solver options and tolerance/temperature
options against the C oracle in the same configuration:

* dense = 1, iterative = 0, analytic = 0: CVODE's dense direct solver with
  its difference-quotient Jacobian (cvLsDenseDQJac);
* SPGMR (dense = 0 or iterative = 1; solver_cvode.cpp:186-214) preconditioned
  by I - gamma J with the threshold of FormPreconditioner
  (reactor_nvector_serial.cpp:414-493): analytic J with and without the
  threshold (iterative = 0 keeps every entry of the pattern), the reactor's
  divided differences (analytic = 0, :816-835) on the Jacobian's pattern,
  and the full divided-difference matrix (dense = 1, analytic = 0);
* abstol_dens (reactor_nvector_serial.cpp:107-120);
* always_solve_temperature = 0 through the plugin (the temperature held
  fixed, zerork_reactor_manager.cpp:1109-1113) and the reactor timing log.

The bar is the one of test_gpu_integrate.py: the end-state differences
(tolerance units rtol |y| + atol) within the oracle's own one-ulp
sensitivity in the same mode, the mean step count within 2 % (5 % with
FP32 matrices) plus that sensitivity, no failures.  Through the plugin, the
reference tester's scenario with the reference's SPGMR configuration also
takes the reference's own logged number of CVode calls per solve
(examples/cfd_plugin_tester/outputs.ref/*_sparse.log), as the oracle does
(test_oracle_tester.py).
"""
import os

import numpy as np
import pytest

from conftest import mech_paths
from test_gpu_integrate import _gpu_integrate, _units
from test_gpu_team import _dme_states
from test_oracle_tester import (IDT, TESTER, THRESHOLD, initial_mass_fractions,
                                reference_counts)

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-6, 1.0e-12
MODES = {
    "direct_dq": dict(dense=1, analytic=0, iterative=0),
    "spgmr": dict(dense=0, analytic=1, iterative=1),
    "spgmr_full": dict(dense=0, analytic=1, iterative=0),
    "spgmr_reactor_dq": dict(dense=0, analytic=0, iterative=1),
    "spgmr_dense_dq": dict(dense=1, analytic=0, iterative=1),
}


def _abstol_dens(o, T, P, Y, atol):
    """The species absolute tolerances that abstol_dens = 1 gives each reactor:
    atol / (rho0 / MW_j 1e-3), rho0 the initial density
    (reactor_nvector_serial.cpp:107-120)."""
    mw = o.tables.mw
    rho0 = P / (8.314462618e3 * T) / np.sum(Y / mw, axis=1)
    return atol * mw[None, :] * 1.0e3 / rho0[:, None]


def _compare(Tg, Yg, steps, To, Yo, ost, Tp, Yp, pst, f32, label, atol=ATOL):
    """atol: the absolute tolerance of the run (a scalar, or per reactor and
    species), so that differences are counted in the run's own units."""
    gpu = _units(Yg, Tg, Yo, To, RTOL, atol)
    ulp = _units(Yp, Tp, Yo, To, RTOL, atol)
    q = [50, 90, 100]
    gq, uq = np.percentile(gpu, q), np.percentile(ulp, q)
    print(f"{label}: GPU-oracle tolerance units p50/p90/max {gq.round(2)}; oracle 1-ulp "
          f"{uq.round(2)}; mean steps gpu {steps.mean():.1f} oracle {ost['nsteps'].mean():.1f}",
          flush=True)
    assert np.all(ost["nsteps"] > 0)
    assert np.all(steps > 0), np.sum(steps <= 0)
    slack = (10.0, 50.0) if f32 else (1.0, 10.0)
    assert gq[0] <= 2.0 * uq[0] + 1.0 and gq[1] <= 2.0 * uq[1] + slack[0]
    assert gq[2] <= 4.0 * uq[2] + slack[1]
    dsteps = abs(pst["nsteps"].mean() / ost["nsteps"].mean() - 1.0)
    assert abs(steps.mean() / ost["nsteps"].mean() - 1.0) < (0.05 if f32 else 0.02) + 2.0 * dsteps


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("wl,n,seed,f32", [("10sp", 512, 201, 0), ("53sp", 128, 202, 0),
                                           ("53sp", 128, 202, 1)])
def test_solver_mode_matches_oracle(torch_gpu, oracle_mod, mode, wl, n, seed, f32):
    from zerork_hip import capi, mechgen
    name = f"synth_{wl}"
    m = capi.Mech(*mech_paths(name))
    o = oracle_mod.Oracle(*mech_paths(name))
    mg = mechgen.workload_mechanism(wl)
    st = mechgen.workload_states(wl, mg, n, seed=seed)
    dt = mechgen.WORKLOAD_DT[wl]
    kw = dict(MODES[mode], eps_lin=0.05, preconditioner_threshold=1e-4)
    print("case", mode, wl, "n", n, "fp32", f32, flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, fp32_preconditioner=f32, **kw)
    Tg, Yg, steps, stats = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, dt, opts)
    oo = oracle_mod.default_opts(rtol=RTOL, atol=ATOL, **kw)
    To, Yo, ost = o.integrate(st.T, st.P, st.Y, dt, oo)
    Tp, Yp, pst = o.integrate(np.nextafter(st.T, np.inf), st.P, st.Y, dt, oo)
    _compare(Tg, Yg, steps, To, Yo, ost, Tp, Yp, pst, f32, f"{mode} {wl} fp32={f32}")


@pytest.mark.parametrize("mode", ["direct", "spgmr"])
def test_teams_solver_modes(torch_gpu, oracle_mod, mode):
    """DME (79 species): two-wavefront teams, FP64 Newton matrices (the plain
    method's bar: the modes' parity, not the FP32 preconditioner's, is what
    is tested here)."""
    from zerork_hip import capi, mechgen
    m = capi.Mech(*mech_paths("ref_dme"))
    o = oracle_mod.Oracle(*mech_paths("ref_dme"))
    T0, P0, Y0 = _dme_states(m.species(), 64, seed=203)
    st = mechgen.States(T0, P0, Y0, 203)
    dt = 1e-4
    kw = dict(MODES["spgmr"] if mode == "spgmr" else {}, eps_lin=0.05,
              preconditioner_threshold=1e-4, abstol_dens=1)
    print("case dme", mode, flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, fp32_preconditioner=0, **kw)
    Tg, Yg, steps, stats = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, dt, opts)
    oo = oracle_mod.default_opts(rtol=RTOL, atol=ATOL, **kw)
    To, Yo, ost = o.integrate(st.T, st.P, st.Y, dt, oo)
    Tp, Yp, pst = o.integrate(np.nextafter(st.T, np.inf), st.P, st.Y, dt, oo)
    _compare(Tg, Yg, steps, To, Yo, ost, Tp, Yp, pst, 0, f"dme {mode}",
             atol=_abstol_dens(o, st.T, st.P, st.Y, ATOL))


@pytest.mark.parametrize("cvol", [0, 1])
def test_abstol_dens(torch_gpu, oracle_mod, cvol):
    """abstol_dens = 1: species tolerances abs_tol MW_j 1e3 / rho0; with a
    molar-concentration tolerance of 1e-12 kmol/m^3 (1e-9 mol/cm^3) they are
    ~1e-9..1e-6 in mass fraction, far looser than abs_tol: the step counts
    drop, and the GPU follows the oracle."""
    from zerork_hip import capi, mechgen
    m = capi.Mech(*mech_paths("synth_10sp"))
    o = oracle_mod.Oracle(*mech_paths("synth_10sp"))
    mg = mechgen.workload_mechanism("10sp")
    st = mechgen.workload_states("10sp", mg, 512, seed=204)
    print("case abstol_dens cvol", cvol, flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=1e-12, constant_volume=cvol, abstol_dens=1)
    Tg, Yg, steps, _ = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, 1e-4, opts)
    oo = oracle_mod.default_opts(rtol=RTOL, atol=1e-12, constant_volume=cvol, abstol_dens=1)
    To, Yo, ost = o.integrate(st.T, st.P, st.Y, 1e-4, oo)
    Tp, Yp, pst = o.integrate(np.nextafter(st.T, np.inf), st.P, st.Y, 1e-4, oo)
    plain = o.integrate(st.T, st.P, st.Y, 1e-4,
                        oracle_mod.default_opts(rtol=RTOL, atol=1e-12, constant_volume=cvol))
    assert ost["nsteps"].mean() < plain[2]["nsteps"].mean()
    _compare(Tg, Yg, steps, To, Yo, ost, Tp, Yp, pst, 0, f"abstol_dens cvol={cvol}",
             atol=_abstol_dens(o, st.T, st.P, st.Y, 1e-12))


def _tester_plugin(capi, name, opts):
    mech = (os.path.join(TESTER, f"{name}.mech"), os.path.join(TESTER, f"{name}.therm"))
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech) == 0
    for k, v in opts.items():
        assert (p.set_int(k, v) if isinstance(v, int) else p.set_double(k, v)) == 0, k
    assert p.load_mechanism() == 0
    return p, mech


@pytest.mark.parametrize("name", ["h2", "dme"])
def test_plugin_spgmr_takes_reference_steps(torch_gpu, oracle_mod, tmp_path, name):
    """The tester scenario through zerork_reactor_solve with the reference's
    SPGMR configuration (dense 0, analytic 1, iterative 1, eps_lin 0.1, the
    run's preconditioner_threshold): the COST of every solve within max(4,
    12 %) of the reference's logged count and the total within 5 % (the
    oracle's own bar), the states within the tester bound of the oracle's
    trajectory in the same mode, and the timing log written per solve."""
    from zerork_hip import capi
    log = tmp_path / "timing.log"
    p, mech = _tester_plugin(capi, name, {
        "rel_tol": 1e-8, "abs_tol": 1e-18, "max_steps": 3000, "cvode_num_retries": 3,
        "delta_temperature_ignition": 400.0, "dense": 0, "analytic": 1, "iterative": 1,
        "eps_lin": 0.1, "preconditioner_threshold": THRESHOLD[name]})
    assert p.set_string("reactor_timing_log_filename", str(log)) == 0
    o = oracle_mod.Oracle(*mech)
    Y0 = initial_mass_fractions(o, name, mech)
    dt = 2.0 * IDT[name] / 10.0
    oo = oracle_mod.default_opts(rtol=1e-8, atol=1e-18, constant_volume=1, max_steps=3000,
                                 num_retries=3, delta_temperature_ignition=400.0, dense=0,
                                 analytic=1, iterative=1, eps_lin=0.1,
                                 preconditioner_threshold=THRESHOLD[name])
    print("case plugin spgmr", name, flush=True)
    n = 32
    T, P, Y = np.full(n, 1450.0), np.full(n, 2.02650e6), np.tile(Y0, (n, 1))
    cost = np.zeros(n)
    assert p.set_aux("COST", cost) == 0
    To, Po, Yo = T[:1].copy(), P[:1].copy(), Y[:1].copy()
    counts, worst = [], 0.0
    for s in range(10):
        assert p.solve(s, s * dt, dt, T, P, Y) == 0
        assert np.all(cost == cost[0]) and np.all(T == T[0])
        counts.append(int(cost[0]))
        To, Yo, ost = o.integrate(To, Po, Yo, dt, oo)
        Po = ost["P"]
        v = np.concatenate([[T[0], P[0]], Y[0]])
        vo = np.concatenate([To, Po, Yo[0]])
        worst = max(worst, float(np.max(np.abs(v - vo) / (1e-5 * np.abs(vo) + 1e-14))))
    ref = reference_counts(f"{name}_sparse")
    got = np.array(counts)
    print(f"{name}: gpu COST {got.tolist()} reference {ref.tolist()}; worst state "
          f"difference / (1e-5 |x| + 1e-14) {worst:.3f}", flush=True)
    assert np.all(np.abs(got - ref) <= np.maximum(4, 0.12 * ref)), (got, ref)
    assert abs(got.sum() - ref.sum()) <= 0.05 * ref.sum()
    assert worst <= 1.0
    p.free()
    rows = [l.split() for l in open(log) if not l.startswith("#")]
    assert len(rows) == 10 and open(log).readline().startswith("#")
    assert [int(r[0]) for r in rows] == list(range(10))
    assert [int(float(r[7])) for r in rows] == counts  # n_steps_avg


def test_plugin_fixed_temperature(torch_gpu, oracle_mod):
    """always_solve_temperature = 0 with no TEMPERATURE_DELTA field: the
    manager's own temperature changes start at 0, so every reactor is solved
    at its initial temperature (T returned bitwise unchanged) with the
    species equations only, as the oracle's solve_temperature = 0; with the
    field set above 0 for some reactors those are solved with temperature."""
    from zerork_hip import capi, mechgen
    mech = mech_paths("synth_10sp")
    mg = mechgen.workload_mechanism("10sp")
    st = mechgen.workload_states("10sp", mg, 256, seed=205)
    o = oracle_mod.Oracle(*mech)
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech) == 0
    for k, v in (("rel_tol", RTOL), ("abs_tol", ATOL), ("constant_volume", 0),
                 ("always_solve_temperature", 0), ("dense", 1), ("iterative", 0)):
        assert (p.set_int(k, v) if isinstance(v, int) else p.set_double(k, v)) == 0, k
    assert p.load_mechanism() == 0
    print("case plugin fixed temperature", flush=True)
    T, P, Y = st.T.copy(), st.P.copy(), st.Y.copy()
    cost = np.zeros(len(T))
    assert p.set_aux("COST", cost) == 0
    assert p.solve(0, 0.0, 1e-4, T, P, Y) == 0
    np.testing.assert_array_equal(T, st.T)
    oo = oracle_mod.default_opts(rtol=RTOL, atol=ATOL, solve_temperature=0)
    To, Yo, ost = o.integrate(st.T, st.P, st.Y, 1e-4, oo)
    Tp, Yp, pst = o.integrate(st.T, st.P * (1 + 2.2e-16), st.Y, 1e-4, oo)
    np.testing.assert_array_equal(To, st.T)
    steps = cost.astype(int)
    _compare(T, Y, steps, To, Yo, ost, Tp, Yp, pst, 0, "fixed temperature")
    # reactors with a previous temperature change are solved with temperature
    td = np.where(np.arange(len(T)) % 2 == 0, 5.0, 0.0)
    assert p.set_aux("TEMPERATURE_DELTA", td) == 0
    T2, P2, Y2 = st.T.copy(), st.P.copy(), st.Y.copy()
    assert p.solve(1, 0.0, 1e-4, T2, P2, Y2) == 0
    np.testing.assert_array_equal(T2[1::2], st.T[1::2])
    To2, Yo2, _ = o.integrate(st.T[0::2], st.P[0::2], st.Y[0::2], 1e-4,
                              oracle_mod.default_opts(rtol=RTOL, atol=ATOL))
    u = _units(Y2[0::2], T2[0::2], Yo2, To2, RTOL, ATOL)
    assert np.median(u) <= 5.0 and np.any(np.abs(T2[0::2] - st.T[0::2]) > 1e-3)
    p.free()
