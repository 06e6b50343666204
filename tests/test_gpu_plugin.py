"""The drop-in boundary end to end on the GPU: zerork_reactor_solve (host arrays
in, host arrays updated in place) as zerork_cfd_plugin_tester.cpp:333-375 drives
it, against the device-level integrator it wraps and the plugin's output
fields (zerork_reactor_manager.cpp:1110-1139): COST = CVODE steps,
TEMPERATURE_DELTA = T rise above solve_temperature_threshold else 0, GPU =
seconds per reactor, FAILED_SOLVE with the input state kept on failure.
"""
import numpy as np
import pytest

from conftest import mech_paths, plugin_solver_options

pytestmark = pytest.mark.gpu


def _plugin(capi, name, **opts):
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech_paths(name)) == 0
    for k, v in opts.items():
        rc = p.set_int(k, v) if isinstance(v, int) else p.set_double(k, v)
        assert rc == 0, k
    assert p.load_mechanism() == 0
    return p


@pytest.mark.parametrize("cvol", [0, 1])
def test_plugin_solve_matches_device_path(torch_gpu, cvol):
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    mg = mechgen.workload_mechanism("10sp")
    st = mechgen.workload_states("10sp", mg, 256, seed=31)
    n, nsp = st.Y.shape
    print("case plugin 10sp constant_volume", cvol, flush=True)
    p = _plugin(capi, "synth_10sp", constant_volume=cvol, rel_tol=1e-6, abs_tol=1e-12)
    T, P, Y = st.T.copy(), st.P.copy(), st.Y.copy()
    cost = np.zeros(n)
    gpu = np.zeros(n)
    tdelta = np.zeros(n)
    for f, a in (("COST", cost), ("GPU", gpu), ("TEMPERATURE_DELTA", tdelta)):
        assert p.set_aux(f, a) == 0
    assert p.solve(0, 0.0, 1e-4, T, P, Y) == 0
    # the same batch through the device-level entry point
    m = capi.Mech(*mech_paths("synth_10sp"))
    dev = torch.device("cuda:0")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    dT, dP, dY = t(st.T), t(st.P), t(st.Y)
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    m.integrate(n, 1e-4, dT, dP, dY, nsp,
                capi.default_options(rel_tol=1e-6, abs_tol=1e-12, constant_volume=cvol,
                                     **plugin_solver_options(p)), steps)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(T, dT.cpu().numpy())
    np.testing.assert_array_equal(Y, dY.cpu().numpy())
    np.testing.assert_array_equal(P, dP.cpu().numpy())
    if not cvol:
        np.testing.assert_array_equal(P, st.P)
    np.testing.assert_array_equal(cost, steps.cpu().numpy())
    assert np.all(cost > 0) and np.all(gpu > 0)
    rise = T - st.T
    np.testing.assert_array_equal(tdelta, np.where(rise < 2.0, 0.0, rise))


def test_plugin_failed_solve_keeps_state(torch_gpu):
    from zerork_hip import capi, mechgen
    mg = mechgen.workload_mechanism("10sp")
    st = mechgen.workload_states("10sp", mg, 32, seed=32)
    print("case plugin forced failure", flush=True)
    p = _plugin(capi, "synth_10sp", constant_volume=0, rel_tol=1e-6, abs_tol=1e-12,
                max_steps=1, cvode_num_retries=2)
    T, P, Y = st.T.copy(), st.P.copy(), st.Y.copy()
    cost = np.zeros(len(T))
    assert p.set_aux("COST", cost) == 0
    assert p.solve(0, 0.0, 1e-4, T, P, Y) == 7  # ZERORK_STATUS_FAILED_SOLVE
    np.testing.assert_array_equal(T, st.T)
    np.testing.assert_array_equal(Y, st.Y)
    assert np.all(cost < 0)


def test_plugin_reactor_dumps(torch_gpu, tmp_path, monkeypatch):
    """dump_reactors (zerork_reactor_manager.cpp:904-906, 1042-1045, 1477-1497):
    per call, dumpfile_pre_<call>_<rank> and dumpfile_postg_<call>_<rank> with
    one line per reactor "id T P cost gpu mf..." in the solve order (descending
    previous cost), at 20 significant digits (exact round trip)."""
    from zerork_hip import capi, mechgen
    monkeypatch.chdir(tmp_path)
    st = mechgen.workload_states("10sp", mechgen.workload_mechanism("10sp"), 64, seed=44)
    n, nsp = st.Y.shape
    print("case plugin dump_reactors", flush=True)
    p = _plugin(capi, "synth_10sp", constant_volume=0, rel_tol=1e-6, abs_tol=1e-12,
                dump_reactors=1)
    T, P, Y = st.T.copy(), st.P.copy(), st.Y.copy()
    cost = np.zeros(n)
    gpu = np.zeros(n)
    assert p.set_aux("COST", cost) == 0 and p.set_aux("GPU", gpu) == 0
    for call in (1, 2):
        T0, P0, Y0, c0 = T.copy(), P.copy(), Y.copy(), cost.copy()
        assert p.solve(call - 1, 0.0, 1e-5, T, P, Y) == 0
        order = np.argsort(-c0, kind="stable")
        for tag, (Tx, Px, Yx, cx) in (("pre", (T0, P0, Y0, c0)), ("postg", (T, P, Y, cost))):
            rows = [l.rstrip("\n").rstrip("\t").split("\t")
                    for l in open(tmp_path / f"dumpfile_{tag}_{call:06d}_000000")]
            assert len(rows) == n
            a = np.array([[float(v) for v in r] for r in rows])
            np.testing.assert_array_equal(a[:, 0], np.arange(n))
            np.testing.assert_array_equal(a[:, 1], Tx[order])
            np.testing.assert_array_equal(a[:, 2], Px[order])
            np.testing.assert_array_equal(a[:, 3], cx[order])
            np.testing.assert_array_equal(a[:, 5:], Yx[order])
