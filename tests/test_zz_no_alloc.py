"""zrk_integrate_device_noalloc (the integrator with its scratch in a static
device array, for callers that cannot allocate) against zrk_integrate_device,
called through the C ABI: two calls in a row, the second with a smaller
batch, must give bit-identical end states, step counts and statistics
(each reactor's integration does not depend on which team slot runs it).
Covers the one-wavefront kernels (FP64 and FP32 Newton matrices, 53 species)
and the three-wavefront team kernel (160 species), then the host-side
refusals: no device image yet (-6) and a mechanism on the sparse path (-4)."""
import ctypes as C

import numpy as np
import pytest

from conftest import mech_paths

pytestmark = pytest.mark.gpu


def _call(fn, m, opts, st, dt, torch):
    dev = torch.device("cuda:0")
    n, nsp = st.Y.shape
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    T, P, Y = t(st.T), t(st.P), t(st.Y)
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    stats = torch.zeros(8 * n, dtype=torch.int32, device=dev)
    rc = fn(m.h, C.byref(opts), n, dt, T.data_ptr(), P.data_ptr(), Y.data_ptr(), nsp,
            steps.data_ptr(), stats.data_ptr(), None)
    torch.cuda.synchronize()
    return rc, T.cpu().numpy(), P.cpu().numpy(), Y.cpu().numpy(), steps.cpu().numpy(), \
        stats.cpu().numpy()


@pytest.mark.parametrize("wl,fp32", [("53sp", 0), ("53sp", 1), ("160sp", 1)])
def test_noalloc_matches_integrate_device(torch_gpu, wl, fp32):
    from zerork_hip import capi, mechgen
    L = capi.lib()
    m = capi.Mech(*mech_paths(f"synth_{wl}"))
    mg = mechgen.workload_mechanism(wl)
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12, fp32_preconditioner=fp32)
    dt = mechgen.WORKLOAD_DT[wl]
    assert m.resident_waves(opts) > 0  # builds the device image
    for call, n in enumerate((600, 90)):
        st = mechgen.workload_states(wl, mg, n, seed=700 + call)
        print(f"case noalloc {wl} fp32={fp32} call {call} n={n}", flush=True)
        ref = _call(L.zrk_integrate_device, m, opts, st, dt, torch_gpu)
        got = _call(L.zrk_integrate_device_noalloc, m, opts, st, dt, torch_gpu)
        assert ref[0] == 0 and got[0] == 0, (ref[0], got[0])
        assert np.all(ref[4] > 0), "reference integration failed"
        for a, b, what in zip(ref[1:], got[1:], ("T", "P", "Y", "nsteps", "stats")):
            np.testing.assert_array_equal(b, a, err_msg=what)


def test_noalloc_refusals(torch_gpu):
    from zerork_hip import capi, mechgen
    L = capi.lib()
    st = mechgen.workload_states("10sp", mechgen.workload_mechanism("10sp"), 4, seed=71)
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12, fp32_preconditioner=1)
    fresh = capi.Mech(*mech_paths("synth_10sp"))
    rc = _call(L.zrk_integrate_device_noalloc, fresh, opts, st, 1e-4, torch_gpu)[0]
    assert rc == -6, rc  # no device image yet: nothing may be allocated here
    assert fresh.resident_waves(opts) > 0
    fresh.set_sparse(True)
    rc = _call(L.zrk_integrate_device_noalloc, fresh, opts, st, 1e-4, torch_gpu)[0]
    assert rc == -4, rc
