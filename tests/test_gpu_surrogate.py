"""GPU checks of the learned-surrogate kernels (include/zrk_surrogate.h).

* The BF16 MFMA GEMM (any strides, sizes not multiples of 16) against numpy on
  the same BF16-rounded operands (FP64 accumulation): this pins the MFMA
  operand/result lane layout.
* Forward pass and weight/bias gradients of a small network against the CPU
  reference oracle/mlp_ref.py (same BF16 rounding of GEMM operands).
* Training on the GPU reduces the loss on a smooth regression.
* The acceptance test against a numpy restatement of its formula on the
  oracle's RHS, and the per-reactor-dt subset integration against the
  plain integrator.
Tolerances: GEMM 1e-5 relative (FP32 vs FP64 accumulation of exact BF16
products); network 2e-2 relative in norm (an activation near a BF16 rounding
boundary may round the other way on one side).
"""
import numpy as np
import pytest

from conftest import mech_paths

pytestmark = pytest.mark.gpu


def _bf16_tensor(torch, a):
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).cuda()


@pytest.mark.parametrize("M,N,K,trA,trB", [(16, 16, 16, 0, 0), (37, 70, 53, 0, 0),
                                            (64, 48, 100, 1, 0), (33, 17, 40, 0, 1),
                                            (130, 64, 256, 1, 1)])
def test_gemm_bf16_matches_numpy(torch_gpu, M, N, K, trA, trB):
    from zerork_hip import surrogate
    from oracle import mlp_ref
    torch = torch_gpu
    rng = np.random.default_rng(M * N + K)
    A = mlp_ref.bf16(rng.standard_normal((M, K)).astype(np.float32))
    B = mlp_ref.bf16(rng.standard_normal((K, N)).astype(np.float32))
    print("case gemm", M, N, K, trA, trB, flush=True)
    At = _bf16_tensor(torch, A.T.copy() if trA else A)
    Bt = _bf16_tensor(torch, B.T.copy() if trB else B)
    sA = (1, M) if trA else (K, 1)
    sB = (1, K) if trB else (N, 1)
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda:0")
    surrogate.gemm(At, sA, Bt, sB, C, M, N, K)
    torch.cuda.synchronize()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    got = C.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref)) + 1e-6
    # accumulate mode adds to C
    surrogate.gemm(At, sA, Bt, sB, C, M, N, K, accumulate=True)
    torch.cuda.synchronize()
    assert np.max(np.abs(C.cpu().numpy() - 2 * ref)) <= 2e-5 * np.max(np.abs(ref)) + 1e-6


def test_network_gradients_match_oracle(torch_gpu):
    from zerork_hip import capi, surrogate
    from oracle import mlp_ref
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_10sp"))   # nsp = 10: kin = kout = 16
    s = surrogate.Surrogate(torch, m, hidden=48, layers=2, seed=3)
    rng = np.random.default_rng(5)
    B = 96
    X = mlp_ref.bf16(rng.standard_normal((B, s.kin)).astype(np.float32))
    tgt = rng.standard_normal((B, s.nout)) * 0.5
    for i, b in enumerate(s.net.b):
        b.copy_(torch.as_tensor(rng.standard_normal(b.numel()).astype(np.float32) * 0.1))
    print("case network gradients", flush=True)
    Xt = _bf16_tensor(torch, X)
    out, cache = s.forward(Xt, cache=True)
    dout = torch.zeros_like(out)
    dout[:, :s.nout] = (out[:, :s.nout] - torch.as_tensor(tgt, dtype=torch.float32,
                                                          device="cuda:0")) * (2.0 / (B * s.nout))
    dW, db = s.backward(dout, cache)
    torch.cuda.synchronize()
    Ws = [w.cpu().numpy() for w in s.net.W]
    bs = [b.cpu().numpy() for b in s.net.b]
    ref_out, ref_cache = mlp_ref.forward(X, Ws, bs)
    rel = lambda a, b: np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
    assert rel(out.cpu().numpy(), ref_out) < 2e-2
    g = mlp_ref.mse_grad(ref_out, tgt, s.nout)
    rdW, rdb = mlp_ref.backward(g, Ws, ref_cache)
    for i in range(len(Ws)):
        assert rel(dW[i].cpu().numpy(), rdW[i]) < 2e-2, i
        assert rel(db[i].cpu().numpy(), rdb[i]) < 2e-2, i


def test_training_reduces_loss(torch_gpu):
    from zerork_hip import capi, surrogate
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_10sp"))
    s = surrogate.Surrogate(torch, m, hidden=64, layers=2, seed=1)
    rng = np.random.default_rng(2)
    n = 8192
    x = rng.uniform(-1, 1, (n, s.kin)).astype(np.float32)
    x[:, s.nfeat:] = 0.0
    # smooth targets of the inputs
    tgt = np.zeros((n, s.kout), np.float32)
    tgt[:, :s.nout] = np.tanh(x[:, :s.nout] * 1.5) + 0.3 * x[:, 1:s.nout + 1] ** 2
    X = _bf16_tensor(torch, x)
    T = torch.as_tensor(tgt).cuda()
    print("case training", flush=True)
    losses = s.train(X, T, epochs=6, batch=1024, lr=3e-3)
    print("losses", np.round(losses, 5), flush=True)
    assert losses[-1] < 0.2 * losses[0]


def test_accept_and_subset_integration(torch_gpu, oracle_mod):
    """The acceptance norms equal a numpy restatement of the formula on the
    oracle's RHS, and the per-reactor-dt subset integration reproduces the
    plain integrator reactor by reactor."""
    from zerork_hip import capi, mechgen, surrogate
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_10sp"))
    o = oracle_mod.Oracle(*mech_paths("synth_10sp"))
    mg = mechgen.workload_mechanism("10sp")
    st = mechgen.workload_states("10sp", mg, 64, seed=41)
    n, nsp = st.Y.shape
    rng = np.random.default_rng(4)
    dt = np.exp(rng.uniform(np.log(1e-7), np.log(1e-5), n))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12)
    s = surrogate.Surrogate(torch, m, hidden=32, layers=1)
    print("case accept + subset", flush=True)
    # proposal: the plain integrator's own end state (per-reactor dt)
    T1, P1, Y1 = t(st.T), t(st.P), t(st.Y)
    sel = torch.arange(n, dtype=torch.int32, device="cuda:0")
    steps = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    surrogate._chk(surrogate._lib().zrk_integrate_device_select(
        m.h, __import__("ctypes").byref(opts), n, sel.data_ptr(), t(dt).data_ptr(), T1.data_ptr(),
        P1.data_ptr(), Y1.data_ptr(), nsp, steps.data_ptr(), None, 0), "select")
    torch.cuda.synchronize()
    assert np.all(steps.cpu().numpy() > 0)
    # reactor by reactor with the scalar-dt entry point
    for i in (0, 17, 63):
        Ti, Pi, Yi = t(st.T[i:i + 1]), t(st.P[i:i + 1]), t(st.Y[i:i + 1])
        m.integrate(1, float(dt[i]), Ti, Pi, Yi, nsp, opts)
        torch.cuda.synchronize()
        assert Ti.item() == T1[i].item()
        np.testing.assert_array_equal(Yi.cpu().numpy()[0], Y1[i].cpu().numpy())
    acc, res = s.accept(t(st.T), t(st.P), t(st.Y), T1, Y1, t(dt), opts)
    torch.cuda.synchronize()
    res = res.cpu().numpy()
    T1h, Y1h = T1.cpu().numpy(), Y1.cpu().numpy()
    for i in range(n):
        y0 = np.append(st.Y[i], st.T[i])
        y1 = np.append(Y1h[i], T1h[i])
        f0, f1 = o.rhs(st.P[i], y0)[0], o.rhs(st.P[i], y1)[0]
        w = 1.0 / (1e-6 * np.abs(y1) + 1e-12)
        r = np.sqrt(np.mean(((y1 - y0 - 0.5 * dt[i] * (f0 + f1)) * w) ** 2))
        e = np.sqrt(np.mean((0.5 * dt[i] * (f1 - f0) * w) ** 2))
        assert abs(res[i, 0] - r) <= 1e-6 * r + 1e-9
        assert abs(res[i, 1] - e) <= 1e-6 * e + 1e-9
        assert acc[i].item() == int(r <= 1.0 and e <= 1.0)


def _train_surrogate(torch, capi, surrogate, mechgen, m, n_train, seed, epochs, log=print):
    """Training set from the GPU integrator on the 53-species workload states
    with dt ~ logU[1e-6, 1e-4]; returns a trained Surrogate."""
    mg = mechgen.workload_mechanism("53sp")
    st = mechgen.workload_states("53sp", mg, n_train, seed=seed)
    rng = np.random.default_rng(seed)
    dt = np.exp(rng.uniform(np.log(1e-6), np.log(1e-4), n_train))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    T, P, Y, D = t(st.T), t(st.P), t(st.Y), t(dt)
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12)
    sel = torch.arange(n_train, dtype=torch.int32, device="cuda:0")
    surrogate._chk(surrogate._lib().zrk_integrate_device_select(
        m.h, __import__("ctypes").byref(opts), n_train, sel.data_ptr(), D.data_ptr(),
        T.data_ptr(), P.data_ptr(), Y.data_ptr(), m.nsp, None, None, 0), "select")
    torch.cuda.synchronize()
    T1, Y1 = T.cpu().numpy(), Y.cpu().numpy()
    s = surrogate.Surrogate(torch, m, hidden=256, layers=3, seed=seed)
    s.fit_normalisation(st.T, st.P, st.Y, dt, T1, Y1)
    X = s.features(t(st.T), t(st.P), t(st.Y), D)
    tgt = s.target_matrix(st.T, st.Y, T1, Y1)
    losses = s.train(X, torch.as_tensor(tgt).cuda(), epochs=epochs, batch=2048, lr=1e-3, log=log)
    return s, losses


def test_surrogate_step_53sp(torch_gpu):
    """Surrogate-accelerated step on the 53-species workload: trained on the
    GPU integrator's own results, proposals refined by two FP64 Newton
    iterations and accepted only through the FP64 residual/error test, the
    rest integrated by the plain BDF.  Accepted end states must agree with
    the plain BDF's to the integrator's own scale of differences; the
    acceptance rate with the network and with the network-free proposal
    (y1 = y0, same test) are reported."""
    from zerork_hip import capi, mechgen, surrogate
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_53sp"))
    print("case surrogate 53sp", flush=True)
    s, losses = _train_surrogate(torch, capi, surrogate, mechgen, m, 65536, 11, epochs=12)
    assert losses[-1] < losses[0]
    n = 8192
    st = mechgen.workload_states("53sp", mechgen.workload_mechanism("53sp"), n, seed=12)
    rng = np.random.default_rng(12)
    dt = np.exp(rng.uniform(np.log(1e-6), np.log(1e-4), n))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12)
    # plain BDF on every reactor
    Tp, Pp, Yp = t(st.T), t(st.P), t(st.Y)
    sel = torch.arange(n, dtype=torch.int32, device="cuda:0")
    surrogate._chk(surrogate._lib().zrk_integrate_device_select(
        m.h, __import__("ctypes").byref(opts), n, sel.data_ptr(), t(dt).data_ptr(), Tp.data_ptr(),
        Pp.data_ptr(), Yp.data_ptr(), m.nsp, None, None, 0), "select")
    torch.cuda.synchronize()
    out = {}
    for prop in ("mlp", "hold"):
        T, P, Y = t(st.T), t(st.P), t(st.Y)
        acc, steps = s.solve(T, P, Y, t(dt), opts, proposal=prop, newton=2)
        torch.cuda.synchronize()
        a = acc.cpu().numpy().astype(bool)
        Tg, Yg, Tr, Yr = T.cpu().numpy(), Y.cpu().numpy(), Tp.cpu().numpy(), Yp.cpu().numpy()
        units = np.maximum(np.max(np.abs(Yg - Yr) / (1e-6 * np.abs(Yr) + 1e-12), axis=1),
                           np.abs(Tg - Tr) / (1e-6 * Tr))
        out[prop] = (a.mean(), units[a])
        print(f"{prop}: accepted {100 * a.mean():.1f} %; accepted vs plain BDF tol units "
              f"p50/p99/max {np.percentile(units[a], [50, 99, 100]).round(2) if a.any() else '-'}",
              flush=True)
        # rejected reactors went through the plain BDF: identical results
        np.testing.assert_array_equal(Yg[~a], Yr[~a])
        if a.any():
            assert np.percentile(units[a], 99) <= 100.0
