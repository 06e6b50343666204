"""GPU parity of the analytic Jacobian kernel (zrk_jacobian_device) against the
C oracle's restatement of reactor_nvector_serial.cpp:610-771.

Species columns and the temperature row: per entry 1e-10 of
max(|J_ij|, 1e-3 * row max, 1e-6 * column max) (the column floor covers entries
whose terms cancel analytically and keep a round-off residue).  The temperature column is a forward finite
difference (relative step sqrt(eps)), so it inherits round-off amplified by
1/sqrt(eps): compared at 1e-6 of the column maximum plus four times the
oracle's own spread of that column over one-ulp changes of the state (the
same finite difference through different round-off).
"""
import numpy as np
import pytest

from conftest import load_golden, mech_paths

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["synth_10sp", "synth_53sp", "ref_h2", "ref_h2_no_falloff",
                                  "ref_dme", "synth_160sp"])
def test_jacobian_matches_oracle(torch_gpu, oracle_mod, name):
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
        Jo = o.jacobian(g["P"][i], np.append(g["Y"][i], g["T"][i]))
        A, B = Jg[i][:, :nsp], Jo[:, :nsp]
        scale = np.maximum(np.maximum(np.abs(B), 1e-3 * np.max(np.abs(B), axis=1, keepdims=True)),
                           1e-6 * np.max(np.abs(B), axis=0, keepdims=True))
        assert np.all(np.abs(A - B) <= 1e-10 * scale + 1e-300), (i, np.max(np.abs(A - B) / (scale + 1e-300)))
        cmax = np.max(np.abs(Jo[:, nsp]))
        y = np.append(g["Y"][i], g["T"][i])
        big = int(np.argmax(g["Y"][i]))
        noise = np.zeros(N)
        for k, to in ((nsp, np.inf), (nsp, -np.inf), (big, np.inf)):
            yp = y.copy()
            yp[k] = np.nextafter(yp[k], to)
            noise = np.maximum(noise, np.abs(o.jacobian(g["P"][i], yp)[:, nsp] - Jo[:, nsp]))
        assert np.all(np.abs(Jg[i][:, nsp] - Jo[:, nsp]) <= 1e-6 * cmax + 4.0 * noise), \
            (i, np.max(np.abs(Jg[i][:, nsp] - Jo[:, nsp]) / cmax), np.max(noise) / cmax)
