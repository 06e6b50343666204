"""GPU parity of the FP64 rate kernel (zrk_rates_device) against the golden
vectors of the reference rate core and against the C oracle's RHS.

Tolerances (FP64): rates of progress and creation/destruction rates 1e-12
relative; net rates 1e-12 of max(|create|, |destroy|); dY/dt and dT/dt 1e-12
of the largest term contributing to them.
"""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, load_golden, mech_paths

pytestmark = pytest.mark.gpu


def _run_rates(torch, m, T, P, Y, limiter=1.0e22):
    dev = torch.device("cuda:0")
    n, nsp = Y.shape
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dT, dP, dY = t(T), t(P), t(Y)
    out = {k: torch.zeros(n * s, dtype=torch.float64, device=dev)
           for k, s in (("ydot", nsp + 1), ("rop", m.nstep), ("create", nsp), ("destroy", nsp),
                        ("net", nsp))}
    m.rates(n, dT, dP, dY, nsp, limiter, out["ydot"], out["rop"], out["create"], out["destroy"],
            out["net"])
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().reshape(n, -1) for k, v in out.items()}


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_rates_match_reference_golden(torch_gpu, name):
    from zerork_hip import capi
    g = load_golden(name)
    m = capi.Mech(*mech_paths(name))
    print("case", name, "states", len(g["T"]), flush=True)
    r = _run_rates(torch_gpu, m, g["T"], g["P"], g["Y"])
    np.testing.assert_allclose(r["rop"], g["rop"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(r["create"], g["create"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(r["destroy"], g["destroy"], rtol=1e-12, atol=0)
    scale = np.maximum(np.abs(g["create"]), np.abs(g["destroy"]))
    assert np.all(np.abs(r["net"] - g["net"]) <= 1e-12 * scale)
    r0 = _run_rates(torch_gpu, m, g["T"], g["P"], g["Y"], limiter=0.0)
    np.testing.assert_allclose(r0["rop"], g["rop_nolimit"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("name", ["synth_10sp", "synth_53sp", "synth_160sp", "ref_h2", "ref_dme"])
def test_rhs_matches_oracle(torch_gpu, oracle_mod, name):
    from zerork_hip import capi
    g = load_golden(name)
    m = capi.Mech(*mech_paths(name))
    o = oracle_mod.Oracle(*mech_paths(name))
    r = _run_rates(torch_gpu, m, g["T"], g["P"], g["Y"])
    mw = o.tables.mw
    for i in range(len(g["T"])):
        yd, flag = o.rhs(g["P"][i], np.append(g["Y"][i], g["T"][i]))
        assert flag == 0
        rho = g["P"][i] / (8.314462618e3 * g["T"][i]) / np.sum(g["Y"][i] / mw)
        sY = np.maximum(g["create"][i], g["destroy"][i]) * mw / rho
        assert np.all(np.abs(r["ydot"][i, :-1] - yd[:-1]) <= 1e-12 * sY + 1e-300)
        cp = np.sum(g["cp_R"][i] * 8.314462618e3 / mw * g["Y"][i])
        sT = np.sum(np.abs(g["h_RT"][i]) * np.maximum(g["create"][i], g["destroy"][i])) \
            * 8.314462618e3 * g["T"][i] / (rho * cp)
        assert abs(r["ydot"][i, -1] - yd[-1]) <= 1e-12 * sT


def test_rates_large_batch_properties(torch_gpu):
    """65,536 reactors of the 53-species workload: the per-reactor results do not
    depend on the batch (bitwise equal to a small batch of the same states)
    and dY/dt sums to zero (mass conservation)."""
    from zerork_hip import capi, mechgen
    m = capi.Mech(*mech_paths("synth_53sp"))
    mg = mechgen.workload_mechanism("53sp")
    st = mechgen.workload_states("53sp", mg, 65536, seed=3)
    r = _run_rates(torch_gpu, m, st.T, st.P, st.Y)
    idx = np.array([0, 17, 4095, 65535])
    rs = _run_rates(torch_gpu, m, st.T[idx], st.P[idx], st.Y[idx])
    for k in rs:
        np.testing.assert_array_equal(rs[k], r[k][idx])
    mw = load_golden("synth_53sp")["mw"]
    s = np.abs(r["ydot"][:, :-1]).sum(axis=1)
    assert np.all(np.abs(r["ydot"][:, :-1].sum(axis=1)) <= 1e-12 * s + 1e-300)
    assert np.all(np.isfinite(r["rop"]))
    del mw
