"""The outer CFD step captured in a HIP graph: restore the batch state, then
integrate it (the work-counter reset is a kernel node, not a memset).  Every
replay, with eager work and synchronisations between replays, must reproduce
the eager integration bit for bit."""
import numpy as np
import pytest

from conftest import mech_paths

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp32", [0, 1])
def test_graph_replays_match_eager(torch_gpu, fp32):
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_10sp"))
    st = mechgen.workload_states("10sp", mechgen.workload_mechanism("10sp"), 1024, seed=61)
    n, nsp = st.Y.shape
    print("case graph replay fp32", fp32, flush=True)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    T0, P0, Y0 = t(st.T), t(st.P), t(st.Y)
    T, P, Y = T0.clone(), P0.clone(), Y0.clone()
    steps = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12, fp32_preconditioner=fp32)
    # eager reference (also allocates the integrator workspace before capture)
    m.integrate(n, 1e-4, T, P, Y, nsp, opts, steps)
    torch.cuda.synchronize()
    Te, Ye, se = T.cpu().numpy(), Y.cpu().numpy(), steps.cpu().numpy()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            T.copy_(T0)
            Y.copy_(Y0)
            m.integrate(n, 1e-4, T, P, Y, nsp, opts, steps,
                        stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ydot = torch.empty(n * (nsp + 1), dtype=torch.float64, device="cuda:0")
    for k in range(3):
        T.fill_(0.0)
        Y.fill_(0.0)
        steps.zero_()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(T.cpu().numpy(), Te)
        np.testing.assert_array_equal(Y.cpu().numpy(), Ye)
        np.testing.assert_array_equal(steps.cpu().numpy(), se)
        # eager work between replays
        m.rates(n, T0, P0, Y0, nsp, 1.0e22, ydot)
        torch.cuda.synchronize()
