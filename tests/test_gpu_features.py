"""GPU parity for the extended reaction kinds (SRI fall-off, PLOG, fractional
stoichiometry, FORD/RORD), which the device evaluates in a separate record
loop and, for the non-integer reactions, with exponent lists, species term
lists and Jacobian terms of their own.

* RHS (zrk_rates_device): dY/dt and dT/dt against the C oracle at 1e-12 of the
  largest contributing term (the oracle is pinned to the reference's rate core
  and to the reference tests' closed forms for these mechanisms); the rates of
  progress themselves are checked against the golden vectors in
  test_gpu_rates.py.
* Jacobian (zrk_jacobian_device): against the oracle, with the tolerances of
  test_gpu_jacobian.py.
* Integrator: the H2 mechanism with every extended kind added, and the
  reference's FORD/RORD propane mechanisms, constant pressure and constant
  volume, FP64 and the FP32 preconditioner: self-calibrated end-state parity
  with the oracle (its own one-ulp sensitivity), step counts, sum(Y), and
  accuracy against the LSODA referee, as in test_gpu_integrate.py.
"""
import numpy as np
import pytest

from conftest import FEATURE_CASES, load_golden, mech_paths
from test_gpu_integrate import _gpu_integrate, _referee, _units
from test_gpu_rates import _run_rates

pytestmark = pytest.mark.gpu

RU = 8.314462618e3
RTOL, ATOL = 1.0e-6, 1.0e-12


@pytest.mark.parametrize("name", FEATURE_CASES)
def test_feature_rhs_matches_oracle(torch_gpu, oracle_mod, name):
    from zerork_hip import capi
    g = load_golden(name)
    m = capi.Mech(*mech_paths(name))
    o = oracle_mod.Oracle(*mech_paths(name))
    print("case", name, "states", len(g["T"]), flush=True)
    r = _run_rates(torch_gpu, m, g["T"], g["P"], g["Y"])
    mw = o.tables.mw
    for i in range(len(g["T"])):
        yd, flag = o.rhs(g["P"][i], np.append(g["Y"][i], g["T"][i]))
        assert flag == 0
        rho = g["P"][i] / (RU * g["T"][i]) / np.sum(g["Y"][i] / mw)
        cd = np.maximum(np.abs(g["create"][i]), np.abs(g["destroy"][i]))
        sY = cd * mw / rho
        assert np.all(np.abs(r["ydot"][i, :-1] - yd[:-1]) <= 1e-12 * sY + 1e-300), i
        cp = np.sum(g["cp_R"][i] * RU / mw * g["Y"][i])
        sT = np.sum(np.abs(g["h_RT"][i]) * cd) * RU * g["T"][i] / (rho * cp)
        assert abs(r["ydot"][i, -1] - yd[-1]) <= 1e-12 * sT, i


@pytest.mark.parametrize("name", FEATURE_CASES)
def test_feature_jacobian_matches_oracle(torch_gpu, oracle_mod, name):
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
    for i in range(n):
        y = np.append(g["Y"][i], g["T"][i])
        Jo = o.jacobian(g["P"][i], y)
        A, B = Jg[i][:, :nsp], Jo[:, :nsp]
        scale = np.maximum(np.maximum(np.abs(B), 1e-3 * np.max(np.abs(B), axis=1, keepdims=True)),
                           1e-6 * np.max(np.abs(B), axis=0, keepdims=True))
        assert np.all(np.abs(A - B) <= 1e-10 * scale + 1e-300), (i, np.max(np.abs(A - B) / (scale + 1e-300)))
        # temperature column: forward difference; spread of the oracle's own
        # column over one-ulp changes of the state bounds the round-off
        cols = [o.jacobian(g["P"][i], np.append(g["Y"][i], np.nextafter(g["T"][i], d)))[:, nsp]
                for d in (np.inf, -np.inf)]
        spread = np.max(np.abs(np.array(cols) - Jo[:, nsp]), axis=0)
        cs = np.max(np.abs(Jo[:, nsp]))
        assert np.all(np.abs(Jg[i][:, nsp] - Jo[:, nsp]) <= 1e-6 * cs + 4.0 * spread), i


def _h2_states(species, n, seed):
    rng = np.random.default_rng(seed)
    sp = {s: j for j, s in enumerate(species)}
    Y = np.full((n, len(species)), 1e-7)
    Y[:, sp["h2"]] = rng.uniform(0.01, 0.05, n)
    Y[:, sp["o2"]] = rng.uniform(0.15, 0.25, n)
    Y[:, sp["h2o"]] = rng.uniform(0.0, 0.05, n)
    Y[:, sp["n2"]] = 0.0
    Y[:, sp["n2"]] = 1.0 - Y.sum(axis=1)
    T = rng.uniform(950.0, 1500.0, n)
    P = np.exp(rng.uniform(np.log(0.5), np.log(30.0), n)) * 101325.0
    return T, P, Y


def _propane_states(species, n, seed):
    rng = np.random.default_rng(seed)
    sp = {s: j for j, s in enumerate(species)}
    Y = np.full((n, len(species)), 1e-4)
    Y[:, sp["c3h8"]] = rng.uniform(0.02, 0.06, n)
    Y[:, sp["o2"]] = rng.uniform(0.18, 0.23, n)
    Y[:, sp["n2"]] = 0.0
    Y[:, sp["n2"]] = 1.0 - Y.sum(axis=1)
    T = rng.uniform(1300.0, 1800.0, n)
    P = np.exp(rng.uniform(np.log(1.0), np.log(20.0), n)) * 101325.0
    return T, P, Y


@pytest.mark.parametrize("name,cvol,f32,team", [
    ("ref_h2_features", 0, 0, 0), ("ref_h2_features", 1, 0, 0), ("ref_h2_features", 0, 1, 0),
    ("ref_h2_features", 0, 1, 2), ("ref_ford_propane", 0, 0, 0), ("ref_rord_propane", 1, 0, 0),
    ("ref_ford_propane", 0, 1, 0)])
def test_feature_integrate_matches_oracle(torch_gpu, oracle_mod, name, cvol, f32, team):
    """team=2: the two-wavefront team kernel forced on this small mechanism
    (the path of 64-127 species), so its rate, species and Jacobian loops run
    the extended records too."""
    from zerork_hip import capi
    m = capi.Mech(*mech_paths(name))
    if team:
        m.set_team_width(1, team)
    o = oracle_mod.Oracle(*mech_paths(name))
    gen = _h2_states if name == "ref_h2_features" else _propane_states
    n, dt = 256, 1.0e-4
    T, P, Y = gen(o.tables.species, n, 5)
    print("case", name, "n", n, "constant volume", cvol, "fp32 preconditioner", f32, "team", team,
          flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, constant_volume=cvol,
                                fp32_preconditioner=f32)
    Tg, Yg, steps, stats, Pg = _gpu_integrate(torch_gpu, m, T, P, Y, dt, opts, True)
    oopts = oracle_mod.default_opts(rtol=RTOL, atol=ATOL, constant_volume=cvol)
    To, Yo, ost = o.integrate(T, P, Y, dt, oopts)
    Tp, Yp, pst = o.integrate(np.nextafter(T, np.inf), P, Y, dt, oopts)
    assert np.all(ost["nsteps"] > 0)
    assert np.all(steps > 0), np.sum(steps <= 0)
    gpu = _units(Yg, Tg, Yo, To, RTOL, ATOL)
    ulp = _units(Yp, Tp, Yo, To, RTOL, ATOL)
    q = [50, 90, 100]
    gq, uq = np.percentile(gpu, q), np.percentile(ulp, q)
    print(f"{name}: GPU-oracle tolerance units p50/p90/max {gq.round(2)}; oracle 1-ulp "
          f"sensitivity {uq.round(2)}; mean steps gpu {steps.mean():.1f} oracle "
          f"{ost['nsteps'].mean():.1f}", flush=True)
    if not f32:
        assert gq[0] <= 2.0 * uq[0] + 1.0 and gq[1] <= 2.0 * uq[1] + 1.0
        assert gq[2] <= 4.0 * uq[2] + 10.0
    else:
        assert gq[0] <= 2.0 * uq[0] + 1.0 and gq[1] <= 2.0 * uq[1] + 10.0
        assert gq[2] <= 4.0 * uq[2] + 50.0
    if not f32:
        dsteps = abs(pst["nsteps"].mean() / ost["nsteps"].mean() - 1.0)
        assert abs(steps.mean() / ost["nsteps"].mean() - 1.0) < 0.02 + 2.0 * dsteps
    else:
        # with FP32 Newton matrices the few reactors whose fuel runs out (the
        # derivative of |C|^0.9 is singular at C = 0; ~5,000 steps each, in the
        # oracle too) are not the same reactors, and they dominate the mean:
        # compare medians
        dmed = abs(np.median(pst["nsteps"]) / np.median(ost["nsteps"]) - 1.0)
        assert abs(np.median(steps) / np.median(ost["nsteps"]) - 1.0) < 0.05 + 2.0 * dmed
        assert abs(steps.mean() / ost["nsteps"].mean() - 1.0) < 0.15
    osum = max(np.abs(Yo.sum(axis=1) - 1.0).max(), np.abs(Yp.sum(axis=1) - 1.0).max())
    assert np.all(np.abs(Yg.sum(axis=1) - 1.0) <= max(1e-8 if f32 else 1e-9, 2.0 * osum))
    if cvol:
        pu = np.abs(Pg - ost["P"]) / (RTOL * ost["P"])
        pp = np.abs(Tp - To) / (RTOL * To)
        assert np.percentile(pu, 90) <= 2.0 * np.percentile(pp, 90) + 1.0
        assert pu.max() <= 4.0 * max(uq[2], 1.0) + 10.0
    else:
        np.testing.assert_array_equal(Pg, P)
    sub = np.linspace(0, n - 1, 6).astype(int)
    eg, eo = [], []
    for i in sub:
        yr, tr = _referee(o, T[i], P[i], Y[i], dt, oopts)
        eg.append(_units(Yg[i:i + 1], Tg[i:i + 1], yr[None], np.array([tr]), RTOL, ATOL)[0])
        eo.append(_units(Yo[i:i + 1], To[i:i + 1], yr[None], np.array([tr]), RTOL, ATOL)[0])
    print(f"{name}: error vs referee (tol units) gpu {np.round(eg, 1)} oracle {np.round(eo, 1)}",
          flush=True)
    assert max(eg) <= 2.0 * max(eo) + 50.0
