"""GPU parity of the multi-wavefront integrator (mechanisms of 64..191 species,
one workgroup of 2 or 3 wavefronts per reactor; an FP32 Newton matrix
inverted in place in LDS by blocked Gauss-Jordan on FP32 MFMA, or an FP64 one
in the team's device slot on FP64 MFMA) against the FP64 oracle: the seeded
160-species / 900-reaction workload mechanism and the reference's 79-species
DME mechanism (data/mechanisms/dme/dme_24_mech.txt).

As for the one-wavefront integrators, the bar is self-calibrated on the same
inputs: GPU-vs-oracle differences in tolerance units (rtol |y| + atol) within
the oracle's own one-ulp sensitivity (the FP64 matrix with the plain method's
margins, the FP32 one with the FP32 preconditioner's), mean step counts
close, no failures, sum(Y) = 1 to 1e-9 (FP64) or 1e-8 (FP32).
"""
import numpy as np
import pytest

from conftest import mech_paths

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-6, 1.0e-12


def _units(Ya, Ta, Yb, Tb):
    eY = np.abs(Ya - Yb) / (RTOL * np.abs(Yb) + ATOL)
    eT = np.abs(Ta - Tb) / (RTOL * np.abs(Tb))
    return np.maximum(eY.max(axis=1), eT)


def _dme_states(species, n, seed):
    rng = np.random.default_rng(seed)
    idx = {s: i for i, s in enumerate(species)}
    Y = np.zeros((n, len(species)))
    phi = rng.uniform(0.5, 2.0, n)
    # DME + air: CH3OCH3 + 3 O2 -> 2 CO2 + 3 H2O (mass ratio 46.07 : 96)
    fuel = phi * 46.07 / 96.0 * 0.233
    Y[:, idx["ch3och3"]] = fuel / (1.0 + fuel)
    Y[:, idx["o2"]] = 0.233 / (1.0 + fuel)
    Y[:, idx["n2"]] = 1.0 - Y[:, idx["ch3och3"]] - Y[:, idx["o2"]]
    T = rng.uniform(900.0, 1400.0, n)
    P = rng.uniform(1.0, 20.0, n) * 101325.0
    return T, P, Y


@pytest.mark.parametrize("name,n,dt,f32", [("synth_160sp", 48, 1e-5, 1), ("ref_dme", 64, 1e-4, 1),
                                          ("synth_160sp", 48, 1e-5, 0), ("ref_dme", 64, 1e-4, 0)])
def test_team_integrator_matches_oracle(torch_gpu, oracle_mod, name, n, dt, f32):
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    m = capi.Mech(*mech_paths(name))
    o = oracle_mod.Oracle(*mech_paths(name))
    if name == "synth_160sp":
        st = mechgen.workload_states("160sp", mechgen.workload_mechanism("160sp"), n, seed=160)
        T0, P0, Y0 = st.T, st.P, st.Y
    else:
        T0, P0, Y0 = _dme_states(m.species(), n, seed=79)
    nsp = m.nsp
    print("case", name, "nsp", nsp, "n", n, "fp32 preconditioner", f32, flush=True)
    dev = torch.device("cuda:0")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dT, dP, dY = t(T0), t(P0), t(Y0)
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, fp32_preconditioner=f32)
    m.integrate(n, dt, dT, dP, dY, nsp, opts, steps)
    torch.cuda.synchronize()
    Tg, Yg, sg = dT.cpu().numpy(), dY.cpu().numpy(), steps.cpu().numpy()
    oo = oracle_mod.default_opts(rtol=RTOL, atol=ATOL)
    To, Yo, ost = o.integrate(T0, P0, Y0, dt, oo)
    Tp, Yp, _ = o.integrate(np.nextafter(T0, np.inf), P0, Y0, dt, oo)
    assert np.all(ost["nsteps"] > 0)
    assert np.all(sg > 0), np.sum(sg <= 0)
    gpu, ulp = _units(Yg, Tg, Yo, To), _units(Yp, Tp, Yo, To)
    gq, uq = np.percentile(gpu, [50, 90, 100]), np.percentile(ulp, [50, 90, 100])
    print(f"{name}: GPU-oracle tolerance units p50/p90/max {gq.round(2)}; oracle 1-ulp "
          f"{uq.round(2)}; mean steps gpu {sg.mean():.1f} oracle {ost['nsteps'].mean():.1f}; "
          f"T rise max {np.max(To - T0):.1f} K", flush=True)
    if f32:
        # the FP32 preconditioner's bar (as for one wavefront): FP32 Newton
        # matrices change the iterates by more than round-off
        assert gq[0] <= 2.0 * uq[0] + 1.0 and gq[1] <= 2.0 * uq[1] + 10.0
        assert gq[2] <= 4.0 * uq[2] + 50.0
        assert abs(sg.mean() / ost["nsteps"].mean() - 1.0) < 0.05
    else:
        # FP64 throughout: the one-wavefront plain bar, except at p90, which
        # over 64 igniting DME reactors sits a few units above twice the
        # one-ulp p90: the GPU's Gauss-Jordan solve differs from the oracle's
        # pivoted LU by round-off at every Newton iteration, while the
        # one-ulp run perturbs the state once, at t = 0
        assert gq[0] <= 2.0 * uq[0] + 1.0 and gq[1] <= 2.0 * uq[1] + 5.0
        assert gq[2] <= 4.0 * uq[2] + 10.0
        assert abs(sg.mean() / ost["nsteps"].mean() - 1.0) < 0.02
    osum = max(np.abs(Yo.sum(axis=1) - 1.0).max(), np.abs(Yp.sum(axis=1) - 1.0).max())
    assert np.all(np.abs(Yg.sum(axis=1) - 1.0) <= max(1e-8 if f32 else 1e-9, 2.0 * osum))


def test_two_wave_team_on_53_species_and_autotune(torch_gpu, oracle_mod):
    """A 53-species mechanism also runs on a team of two wavefronts (FP32): same
    parity bar as the one-wavefront FP32 integrator; the autotuner times every
    team size that fits (1, 2 and 3 wavefronts) on a batch and keeps the fastest."""
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_53sp"))
    o = oracle_mod.Oracle(*mech_paths("synth_53sp"))
    st = mechgen.workload_states("53sp", mechgen.workload_mechanism("53sp"), 256, seed=102)
    n, nsp = st.Y.shape
    print("case 53sp two-wave team", flush=True)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, fp32_preconditioner=1)
    m.set_team_width(1, 2)
    dT, dP, dY = t(st.T), t(st.P), t(st.Y)
    steps = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    m.integrate(n, 1e-4, dT, dP, dY, nsp, opts, steps)
    torch.cuda.synchronize()
    Tg, Yg, sg = dT.cpu().numpy(), dY.cpu().numpy(), steps.cpu().numpy()
    oo = oracle_mod.default_opts(rtol=RTOL, atol=ATOL)
    To, Yo, ost = o.integrate(st.T, st.P, st.Y, 1e-4, oo)
    Tp, Yp, _ = o.integrate(np.nextafter(st.T, np.inf), st.P, st.Y, 1e-4, oo)
    gpu, ulp = _units(Yg, Tg, Yo, To), _units(Yp, Tp, Yo, To)
    gq, uq = np.percentile(gpu, [50, 90, 100]), np.percentile(ulp, [50, 90, 100])
    print(f"53sp team of 2: GPU-oracle p50/p90/max {gq.round(2)}; 1-ulp {uq.round(2)}; "
          f"mean steps {sg.mean():.1f} vs {ost['nsteps'].mean():.1f}", flush=True)
    assert np.all(sg > 0)
    assert gq[0] <= 2.0 * uq[0] + 1.0 and gq[1] <= 2.0 * uq[1] + 10.0
    assert gq[2] <= 4.0 * uq[2] + 50.0
    assert abs(sg.mean() / ost["nsteps"].mean() - 1.0) < 0.05
    # autotune on a larger batch
    m.set_team_width(1, 0)
    big = mechgen.workload_states("53sp", mechgen.workload_mechanism("53sp"), 8192, seed=5)
    choice, ms = m.autotune(8192, 1e-4, t(big.T), t(big.P), t(big.Y), nsp, opts)
    print(f"autotune: chose {choice} wavefronts per reactor; times {ms}", flush=True)
    assert 1 in ms and 2 in ms and set(ms) <= {1, 2, 3}
    assert choice in ms and ms[choice] == min(ms.values())
