"""GPU parity of the persistent BDF integrator (zrk_integrate_device) against the
C oracle's CVODE restatement, on the workload distributions.

Both integrate the same ODE with the same variable-order BDF and options, so
each end state is within the method's global error of the true solution; the
test bound is 100 (rtol |y| + atol) per component (two global errors of up to
~50 tolerance units each, measured for the oracle against an independent
referee in test_oracle_math.py), plus sum(Y) = 1 to 1e-10 and no failures.
"""
import numpy as np
import pytest

from conftest import mech_paths

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-6, 1.0e-12


def _gpu_integrate(torch, m, T, P, Y, dt, opts):
    dev = torch.device("cuda:0")
    n, nsp = Y.shape
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dT, dP, dY = t(T), t(P), t(Y)
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    stats = torch.zeros(n * 8, dtype=torch.int32, device=dev)
    m.integrate(n, dt, dT, dP, dY, nsp, opts, steps, stats)
    torch.cuda.synchronize()
    return (dT.cpu().numpy(), dY.cpu().numpy(), steps.cpu().numpy(),
            stats.cpu().numpy().reshape(n, 8))


def _compare(To, Yo, Tg, Yg, rtol, atol):
    eY = np.abs(Yg - Yo) / (rtol * np.abs(Yo) + atol)
    eT = np.abs(Tg - To) / (rtol * np.abs(To))
    return max(eY.max(), eT.max()), np.median(eY.max(axis=1))


@pytest.mark.parametrize("cfg,n", [(1, 1024), (2, 256)])
def test_integrate_matches_oracle(torch_gpu, oracle_mod, cfg, n):
    from zerork_hip import capi, mechgen
    name = f"synth_cfg{cfg}"
    m = capi.Mech(*mech_paths(name))
    o = oracle_mod.Oracle(*mech_paths(name))
    mg = mechgen.config_mechanism(cfg)
    st = mechgen.config_states(cfg, mg, n, seed=100 + cfg)
    dt = mechgen.CONFIG_DT[cfg]
    print("case cfg", cfg, "n", n, flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL)
    Tg, Yg, steps, stats = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, dt, opts)
    To, Yo, ost = o.integrate(st.T, st.P, st.Y, dt, oracle_mod.default_opts(rtol=RTOL, atol=ATOL))
    assert np.all(ost["nsteps"] > 0)
    assert np.all(steps > 0), np.sum(steps <= 0)
    emax, emed = _compare(To, Yo, Tg, Yg, RTOL, ATOL)
    same_steps = np.mean(steps == ost["nsteps"])
    print(f"cfg{cfg}: max err {emax:.3g} tol units, median {emed:.3g}, "
          f"identical step counts {same_steps:.3f}, mean steps gpu {steps.mean():.1f} "
          f"oracle {ost['nsteps'].mean():.1f}", flush=True)
    assert emax <= 100.0
    assert np.all(np.abs(Yg.sum(axis=1) - 1.0) <= 1e-10)


def test_integrate_full_batch_properties(torch_gpu):
    """The 53-species x 65,536-reactor workload: no failed reactor, sum(Y)=1,
    finite states, and per-reactor results independent of the batch (a
    reactor integrated alone gives bitwise the same end state)."""
    from zerork_hip import capi, mechgen
    m = capi.Mech(*mech_paths("synth_cfg2"))
    mg = mechgen.config_mechanism(2)
    st = mechgen.config_states(2, mg, 65536, seed=2)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL)
    print("case cfg2 full batch", flush=True)
    Tg, Yg, steps, stats = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, 1e-4, opts)
    assert np.all(steps > 0), np.sum(steps <= 0)
    assert np.all(np.isfinite(Yg)) and np.all(np.isfinite(Tg))
    assert np.all(np.abs(Yg.sum(axis=1) - 1.0) <= 1e-10)
    idx = np.array([0, 1, 777, 65535])
    Ts, Ys, ss, _ = _gpu_integrate(torch_gpu, m, st.T[idx], st.P[idx], st.Y[idx], 1e-4, opts)
    np.testing.assert_array_equal(Ts, Tg[idx])
    np.testing.assert_array_equal(Ys, Yg[idx])
    np.testing.assert_array_equal(ss, steps[idx])
    print(f"steps: mean {steps.mean():.1f} max {steps.max()} ; T rise max "
          f"{np.max(Tg - st.T):.1f} K", flush=True)


def test_forced_failure_keeps_state(torch_gpu):
    """max_steps=1 makes every try fail (CV_TOO_MUCH_WORK, solver_cvode.cpp:233-235):
    negative cost, input state unchanged (zerork_reactor_manager.cpp:1123-1139)."""
    from zerork_hip import capi, mechgen
    m = capi.Mech(*mech_paths("synth_cfg1"))
    mg = mechgen.config_mechanism(1)
    st = mechgen.config_states(1, mg, 64, seed=9)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, max_steps=1, cvode_num_retries=3)
    Tg, Yg, steps, stats = _gpu_integrate(torch_gpu, m, st.T, st.P, st.Y, 1e-4, opts)
    assert np.all(steps < 0)
    np.testing.assert_array_equal(Tg, st.T)
    np.testing.assert_array_equal(Yg, st.Y)
