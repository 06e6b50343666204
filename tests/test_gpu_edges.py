"""Edge cases of the device and plugin entry points on the GPU: empty batches
(every path returns success and touches nothing), one reactor, and padded
("ragged") state rows, stride > nsp, whose results must equal the contiguous
layout's bit for bit with the padding left untouched; the rate kernel with
padded rows as well.  The reference plugin accepts n_reactors = 0 and a
caller's stride through its state arrays (zerork_cfd_plugin.cpp:159-175).
"""
import numpy as np
import pytest

from conftest import mech_paths

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-6, 1.0e-12


def _t(torch, a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64,
                           device="cuda:0")


@pytest.mark.parametrize("name,sparse", [("synth_10sp", False), ("synth_53sp", False),
                                         ("synth_160sp", False), ("synth_10sp", True)])
def test_empty_and_single_reactor(torch_gpu, name, sparse):
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    wl = name.split("_")[1]
    m = capi.Mech(*mech_paths(name))
    if sparse:
        m.set_sparse(True)
    nsp = m.nsp
    print("case edges", name, "sparse", sparse, flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL, fp32_preconditioner=1)
    # n = 0: success, nothing written
    T = _t(torch, [1000.0])
    P = _t(torch, [101325.0])
    Y = _t(torch, np.full((1, nsp), 1.0 / nsp))
    T0, Y0 = T.clone(), Y.clone()
    m.integrate(0, 1e-4, T, P, Y, nsp, opts)
    torch.cuda.synchronize()
    assert torch.equal(T, T0) and torch.equal(Y, Y0)
    # one reactor equals the same reactor inside a batch
    st = mechgen.workload_states(wl, mechgen.workload_mechanism(wl), 8, seed=5)
    dT, dP, dY = _t(torch, st.T), _t(torch, st.P), _t(torch, st.Y)
    m.integrate(8, 1e-5, dT, dP, dY, nsp, opts)
    one = [_t(torch, a[3:4]) for a in (st.T, st.P, st.Y)]
    m.integrate(1, 1e-5, one[0], one[1], one[2], nsp, opts)
    torch.cuda.synchronize()
    assert one[0].item() == dT[3].item()
    np.testing.assert_array_equal(one[2].cpu().numpy()[0], dY.cpu().numpy()[3])


@pytest.mark.parametrize("name,pad", [("synth_10sp", 3), ("synth_53sp", 11)])
def test_padded_rows_match_contiguous(torch_gpu, name, pad):
    from zerork_hip import capi, mechgen
    torch = torch_gpu
    wl = name.split("_")[1]
    m = capi.Mech(*mech_paths(name))
    nsp = m.nsp
    n = 96
    st = mechgen.workload_states(wl, mechgen.workload_mechanism(wl), n, seed=9)
    print("case padded rows", name, "stride", nsp + pad, flush=True)
    opts = capi.default_options(rel_tol=RTOL, abs_tol=ATOL)
    # contiguous
    dT, dP, dY = _t(torch, st.T), _t(torch, st.P), _t(torch, st.Y)
    m.integrate(n, 1e-5, dT, dP, dY, nsp, opts)
    # padded: stride nsp + pad, the padding holds a sentinel that must survive
    Yp = np.full((n, nsp + pad), -7.0)
    Yp[:, :nsp] = st.Y
    pT, pP, pY = _t(torch, st.T), _t(torch, st.P), _t(torch, Yp)
    m.integrate(n, 1e-5, pT, pP, pY, nsp + pad, opts)
    torch.cuda.synchronize()
    got = pY.cpu().numpy()
    np.testing.assert_array_equal(pT.cpu().numpy(), dT.cpu().numpy())
    np.testing.assert_array_equal(got[:, :nsp], dY.cpu().numpy())
    assert np.all(got[:, nsp:] == -7.0)
    # the rate kernel reads padded rows the same way
    ydc = torch.zeros(n * (nsp + 1), dtype=torch.float64, device="cuda:0")
    ydp = torch.zeros_like(ydc)
    m.rates(n, _t(torch, st.T), _t(torch, st.P), _t(torch, st.Y), nsp, 1.0e22, ydc)
    m.rates(n, _t(torch, st.T), _t(torch, st.P), _t(torch, Yp), nsp + pad, 1.0e22, ydp)
    torch.cuda.synchronize()
    assert torch.equal(ydc, ydp)


def test_plugin_empty_batch(torch_gpu):
    """zerork_reactor_solve with no reactors returns success."""
    from zerork_hip import capi
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech_paths("synth_10sp")) == 0
    assert p.load_mechanism() == 0
    print("case plugin empty batch", flush=True)
    T, P, Y = np.zeros(0), np.zeros(0), np.zeros(0)
    assert p.solve(0, 0.0, 1e-4, T, P, Y) == 0
