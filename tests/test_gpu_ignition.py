"""GPU parity of the ignition-time root finding (delta_temperature_ignition,
IGNITION_TIME, stop_after_ignition: reactor_nvector_serial.cpp:774-810,
solver_cvode.cpp:90-92, 228-270) against the oracle's restatement of CVODE's
root finding, through zrk_integrate_device_aux and the plugin.

* Root finding does not change the steps: end states bitwise equal to the run
  without it, and exactly one more CVode call per crossing in the cost.
* Ignition times: the same reactors ignite as in the oracle (up to the
  reactors whose crossing the oracle's own one-ulp perturbation moves across
  dt), and the relative time differences stay within the oracle's one-ulp
  sensitivity of its times (quantiles), with a floor of 2e-5 relative (the
  accuracy of a time at rtol 1e-6, 4e-6 against the LSODA referee in the
  oracle tests); FP32 Newton matrices: a floor of 2e-4.
* stop_after_ignition: T = T0 + delta_T at the end to 1e-9 relative, the same
  times as without stopping.
* The plugin fills IGNITION_TIME with the device path's times and counts the
  root returns in COST.
"""
import numpy as np
import pytest

from conftest import mech_paths, plugin_solver_options

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-6, 1.0e-12
DT_IGN = 400.0


def _run(torch, m, st, dt, opts, ign=True):
    dev = torch.device("cuda:0")
    n, nsp = st.Y.shape
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dT, dP, dY = t(st.T), t(st.P), t(st.Y)
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    it = torch.full((n,), -1.0, dtype=torch.float64, device=dev) if ign else None
    m.integrate(n, dt, dT, dP, dY, nsp, opts, steps, ignition_time=it)
    torch.cuda.synchronize()
    return (dT.cpu().numpy(), dY.cpu().numpy(), steps.cpu().numpy(),
            it.cpu().numpy() if ign else None)


@pytest.mark.parametrize("wl,n,cvol,f32", [("10sp", 1024, 0, 0), ("10sp", 512, 1, 0),
                                           ("53sp", 256, 0, 0), ("10sp", 1024, 0, 1)])
def test_ignition_matches_oracle(torch_gpu, oracle_mod, wl, n, cvol, f32):
    from zerork_hip import capi, mechgen
    name = f"synth_{wl}"
    m = capi.Mech(*mech_paths(name))
    o = oracle_mod.Oracle(*mech_paths(name))
    st = mechgen.workload_states(wl, mechgen.workload_mechanism(wl), n, seed=41)
    dt = mechgen.WORKLOAD_DT[wl]
    print("case", wl, n, "constant volume", cvol, "fp32", f32, flush=True)
    base = dict(rel_tol=RTOL, abs_tol=ATOL, constant_volume=cvol, fp32_preconditioner=f32)
    T0, Y0, s0, _ = _run(torch_gpu, m, st, dt, capi.default_options(**base), ign=False)
    T1, Y1, s1, tg = _run(torch_gpu, m, st, dt, capi.default_options(
        delta_temperature_ignition=DT_IGN, **base))
    # root finding leaves the steps alone; one more call per crossing
    np.testing.assert_array_equal(T1, T0)
    np.testing.assert_array_equal(Y1, Y0)
    assert np.all(s0 > 0)
    np.testing.assert_array_equal(s1 - s0, (tg > 0).astype(s1.dtype))
    assert np.all(tg >= 0.0)
    oo = dict(rtol=RTOL, atol=ATOL, constant_volume=cvol, delta_temperature_ignition=DT_IGN)
    _, _, so = o.integrate(st.T, st.P, st.Y, dt, oracle_mod.default_opts(**oo))
    _, _, sp = o.integrate(np.nextafter(st.T, np.inf), st.P, st.Y, dt,
                           oracle_mod.default_opts(**oo))
    to, tp = so["ignition_time"], sp["ignition_time"]
    both = (tg > 0) & (to > 0)
    mism = int(np.sum((tg > 0) != (to > 0)))
    mism_ulp = int(np.sum((tp > 0) != (to > 0)))
    rel = np.abs(tg[both] - to[both]) / to[both]
    bp = (tp > 0) & (to > 0)
    rel_ulp = np.abs(tp[bp] - to[bp]) / to[bp]
    q = [50, 90, 100]
    gq, uq = np.percentile(rel, q), np.percentile(rel_ulp, q)
    print(f"{wl}: ignited gpu {int((tg > 0).sum())} oracle {int((to > 0).sum())}; relative time "
          f"difference p50/p90/max {gq}; oracle 1-ulp {uq}", flush=True)
    assert both.sum() >= 20  # 53 species, dt = 1e-4 s: about 1 reactor in 10 ignites
    assert mism <= 2 * mism_ulp + 2
    floor = 2e-4 if f32 else 2e-5
    assert gq[1] <= 2.0 * uq[1] + floor
    assert gq[2] <= 4.0 * uq[2] + 5.0 * floor


@pytest.mark.parametrize("cvol", [0, 1])
def test_stop_after_ignition(torch_gpu, cvol):
    from zerork_hip import capi, mechgen
    m = capi.Mech(*mech_paths("synth_10sp"))
    st = mechgen.workload_states("10sp", mechgen.workload_mechanism("10sp"), 512, seed=42)
    dt = mechgen.WORKLOAD_DT["10sp"]
    print("case stop_after_ignition constant volume", cvol, flush=True)
    base = dict(rel_tol=RTOL, abs_tol=ATOL, constant_volume=cvol,
                delta_temperature_ignition=DT_IGN)
    _, _, s1, t1 = _run(torch_gpu, m, st, dt, capi.default_options(**base))
    T2, Y2, s2, t2 = _run(torch_gpu, m, st, dt, capi.default_options(stop_after_ignition=1,
                                                                       **base))
    ign = t1 > 0
    assert ign.any()
    np.testing.assert_array_equal(t2, t1)
    Tign = st.T + DT_IGN
    assert np.all(np.abs(T2[ign] - Tign[ign]) <= 1e-9 * Tign[ign])
    assert np.all(s2[ign] <= s1[ign]) and np.all(s2 > 0)
    assert np.all(np.abs(Y2.sum(axis=1) - 1.0) <= 1e-9)


def test_plugin_ignition_time_field(torch_gpu):
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    st = mechgen.workload_states("10sp", mechgen.workload_mechanism("10sp"), 256, seed=43)
    n, nsp = st.Y.shape
    print("case plugin IGNITION_TIME", flush=True)
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech_paths("synth_10sp")) == 0
    for k, v in (("constant_volume", 0), ("rel_tol", 1e-6), ("abs_tol", 1e-12),
                 ("delta_temperature_ignition", DT_IGN)):
        assert (p.set_int(k, v) if isinstance(v, int) else p.set_double(k, v)) == 0, k
    assert p.load_mechanism() == 0
    T, P, Y = st.T.copy(), st.P.copy(), st.Y.copy()
    cost = np.zeros(n)
    ign = np.full(n, -1.0)
    assert p.set_aux("COST", cost) == 0
    assert p.set_aux("IGNITION_TIME", ign) == 0
    assert p.solve(0, 0.0, 1e-4, T, P, Y) == 0
    m = capi.Mech(*mech_paths("synth_10sp"))
    Td, Yd, sd, td = _run(torch, m, st, 1e-4, capi.default_options(
        rel_tol=1e-6, abs_tol=1e-12, delta_temperature_ignition=DT_IGN,
        **plugin_solver_options(p)))
    np.testing.assert_array_equal(ign, td)
    np.testing.assert_array_equal(cost, sd)
    np.testing.assert_array_equal(T, Td)
    assert np.any(ign > 0) and np.all(ign >= 0)
