"""GPU checks of the learned-surrogate kernels (include/zrk_surrogate.h).

* The BF16 MFMA GEMM (any strides, sizes not multiples of 16; the LDS-tiled
  kernel for the network's layouts and the one-tile-per-wavefront kernel for
  the rest) against numpy on the same BF16-rounded operands (FP64
  accumulation): this pins the MFMA operand/result lane layout.
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
                                            (130, 64, 256, 1, 1),
                                            # the LDS-tiled kernel: both layouts of each
                                            # operand, M and K not multiples of the tile
                                            (1000, 256, 64, 0, 0), (256, 64, 4104, 1, 0),
                                            (4096, 64, 256, 0, 1), (512, 128, 96, 1, 1)])
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


@pytest.mark.parametrize("name,hidden,layers", [("synth_10sp", 48, 2), ("synth_53sp", 256, 3)])
def test_network_gradients_match_oracle(torch_gpu, name, hidden, layers):
    """Small network, and the workload's own 56 -> 3 x 256 GELU -> 54 network
    (53 species: kin = 64, kout = 64 after padding)."""
    from zerork_hip import capi, surrogate
    from oracle import mlp_ref
    torch = torch_gpu
    m = capi.Mech(*mech_paths(name))
    s = surrogate.Surrogate(torch, m, hidden=hidden, layers=layers, seed=3)
    rng = np.random.default_rng(5)
    B = 96
    X = mlp_ref.bf16(rng.standard_normal((B, s.kin)).astype(np.float32))
    tgt = rng.standard_normal((B, s.nout)) * 0.5
    for i, b in enumerate(s.net.b):
        b.copy_(torch.as_tensor(rng.standard_normal(b.numel()).astype(np.float32) * 0.1))
    print("case network gradients", name, hidden, layers, flush=True)
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


TAU = (1.0e-3, 4.0e-3, 0.015, 0.05, 0.13, 0.3, 0.6, 1.0)


def _accept_ref(o, P, y0, nodes, dt, tau, rtol, atol, max_newton):
    """numpy restatement of the trajectory acceptance test on the oracle's RHS
    and Jacobian (numpy's LU solve in place of the kernel's Gauss-Jordan
    inverse): (accepted, most iterations, largest first correction, error
    norm where the test stopped, first failed interval)."""
    f = lambda y: o.rhs(P, y)[0]
    N = len(y0)
    t = np.asarray(tau) * dt
    yprev, fprev = y0, f(y0)
    fs = []
    most, first = 0, 0.0
    esum, en = np.zeros(N), 0.0
    for k, yp in enumerate(nodes):
        h = t[k] - (t[k - 1] if k else 0.0)
        c = h if k == 0 else 0.5 * h
        # start point: the proposal, or the previous node if its residual is smaller
        fy = f(yp)
        wp = 1.0 / (rtol * np.abs(yprev) + atol)
        gp = yp - yprev - (h * fy if k == 0 else 0.5 * h * (fprev + fy))
        y = yp.copy() if np.sum((h * fprev * wp) ** 2) >= np.sum((gp * wp) ** 2) else yprev.copy()
        M = np.eye(N) - c * o.jacobian(P, y)
        w = 1.0 / (rtol * np.abs(y) + atol)
        conv, it = False, 0
        while it < max_newton and not conv:
            fy = f(y)
            G = y - yprev - (h * fy if k == 0 else 0.5 * h * (fprev + fy))
            d = np.linalg.solve(M, G)
            y = y - d
            dn = np.sqrt(np.mean((d * w) ** 2))
            if it == 0:
                first = max(first, dn)
            conv = dn <= 0.1
            it += 1
        most = max(most, it)
        if not conv:
            return False, most, first, np.inf, k + 1
        fy = f(y)
        fs.append(fy)
        yprev, fprev = y, fy
        # the local errors that node k completes, weighted at node k
        e = np.zeros(N)
        if k == 1:
            e = np.abs(0.5 * t[0] ** 2 * (fs[1] - fs[0]) / (t[1] - t[0]))
        elif k >= 2:
            a = k - 2
            dd = ((fs[a + 2] - fs[a + 1]) / (t[a + 2] - t[a + 1])
                  - (fs[a + 1] - fs[a]) / (t[a + 1] - t[a])) / (t[a + 2] - t[a])
            e = np.abs((t[k] - t[k - 1]) ** 3 / 12.0 * 2.0 * dd)
            if k == 2:
                e = e + np.abs((t[1] - t[0]) ** 3 / 12.0 * 2.0 * dd)
        esum = esum + e / (rtol * np.abs(y) + atol)
        en = np.sqrt(np.mean(esum ** 2))
        if en > 1.0:
            return False, most, first, en, 0
    return True, most, first, en, 0


def _select(torch, surrogate, m, opts, T, P, Y, dt):
    """zrk_integrate_device_select over every reactor (per-reactor dt), in place."""
    n = T.numel()
    sel = torch.arange(n, dtype=torch.int32, device="cuda:0")
    steps = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    surrogate._chk(surrogate._lib().zrk_integrate_device_select(
        m.h, __import__("ctypes").byref(opts), n, sel.data_ptr(), dt.data_ptr(), T.data_ptr(),
        P.data_ptr(), Y.data_ptr(), m.nsp, steps.data_ptr(), None, 0), "select")
    torch.cuda.synchronize()
    return steps


def test_accept_and_subset_integration(torch_gpu, oracle_mod):
    """The per-reactor-dt subset integration reproduces the plain integrator
    reactor by reactor; the trajectory acceptance test matches its numpy
    restatement on the oracle (flags, failed interval, iterations, norms) for
    three proposals: the plain integrator's own states at the nodes, the same
    states moved by a relative 1e-2, and every node at the initial state.  The
    integrator's states start the Newton iterations with the smallest
    corrections and fail them least often."""
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
    T1, P1, Y1 = t(st.T), t(st.P), t(st.Y)
    steps = _select(torch, surrogate, m, opts, T1, P1, Y1, t(dt))
    assert np.all(steps.cpu().numpy() > 0)
    for i in (0, 17, 63):
        Ti, Pi, Yi = t(st.T[i:i + 1]), t(st.P[i:i + 1]), t(st.Y[i:i + 1])
        m.integrate(1, float(dt[i]), Ti, Pi, Yi, nsp, opts)
        torch.cuda.synchronize()
        assert Ti.item() == T1[i].item()
        np.testing.assert_array_equal(Yi.cpu().numpy()[0], Y1[i].cpu().numpy())
    # the plain integrator's states at the nodes
    Tn, Yn = [], []
    for f in TAU:
        Tk, Pk, Yk = t(st.T), t(st.P), t(st.Y)
        _select(torch, surrogate, m, opts, Tk, Pk, Yk, t(dt * f))
        Tn.append(Tk)
        Yn.append(Yk)
    rates, med = {}, {}
    for label, scale in (("bdf nodes", 0.0), ("perturbed", 1e-2), ("hold", None)):
        Tp = torch.cat(Tn).clone()
        Yp = torch.cat(Yn).clone()
        if scale is None:
            Tp, Yp = t(st.T).repeat(len(TAU)), t(st.Y).repeat(len(TAU), 1)
        elif scale:
            g = torch.Generator(device="cpu").manual_seed(5)
            Yp *= (1.0 + scale * torch.randn(Yp.shape, generator=g, dtype=torch.float64)).cuda()
        Ypin = Yp.cpu().numpy().reshape(len(TAU), n, nsp)
        Tpin = Tp.cpu().numpy().reshape(len(TAU), n)
        acc, res = s.accept(t(st.T), t(st.P), t(st.Y), Tp, Yp, t(dt), opts, TAU, 3)
        torch.cuda.synchronize()
        acc, res = acc.cpu().numpy(), res.cpu().numpy()
        agree = 0
        for i in range(n):
            y0 = np.append(st.Y[i], st.T[i])
            nodes = [np.append(Ypin[k, i], Tpin[k, i]) for k in range(len(TAU))]
            a, most, first, en, fail = _accept_ref(o, st.P[i], y0, nodes, dt[i], TAU, 1e-6,
                                                   1e-12, 3)
            same = a == bool(acc[i]) and fail == res[i, 3]
            agree += int(same)
            if same:  # the same path through the test: the same counts and norms
                assert res[i, 0] == most, (i, res[i], most)
                assert abs(res[i, 1] - first) <= 1e-6 * first + 1e-9, (i, res[i, 1], first)
                if np.isfinite(en):
                    assert abs(res[i, 2] - en) <= 1e-5 * en + 1e-9, (i, res[i, 2], en)
        rates[label] = float(np.mean(res[:, 3] != 0))
        med[label] = float(np.median(res[:, 1]))
        print(f"{label}: accepted {100 * acc.mean():.1f} %, Newton failed {100 * rates[label]:.1f} "
              f"%, median first correction {med[label]:.3g}, agreement with the numpy "
              f"restatement {agree}/{n}", flush=True)
        assert agree >= n - 2
    assert rates["bdf nodes"] <= min(rates["perturbed"], rates["hold"])
    assert med["bdf nodes"] < min(med["perturbed"], med["hold"])


def _train_surrogate(torch, capi, surrogate, mechgen, m, n_train, seed, epochs, log=print):
    """Training set from the GPU integrator on the 53-species workload states
    with step times ~ logU[1e-9, 1e-4] (the proposal nodes tau_k dt of steps
    dt ~ logU[1e-6, 1e-4]); returns a trained Surrogate and its losses."""
    mg = mechgen.workload_mechanism("53sp")
    st = mechgen.workload_states("53sp", mg, n_train, seed=seed)
    rng = np.random.default_rng(seed)
    dt = np.exp(rng.uniform(np.log(1e-9), np.log(1e-4), n_train))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    T, P, Y, D = t(st.T), t(st.P), t(st.Y), t(dt)
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12)
    _select(torch, surrogate, m, opts, T, P, Y, D)
    T1, Y1 = T.cpu().numpy(), Y.cpu().numpy()
    s = surrogate.Surrogate(torch, m, hidden=256, layers=3, seed=seed)
    s.fit_normalisation(st.T, st.P, st.Y, dt, T1, Y1)
    X = s.features(t(st.T), t(st.P), t(st.Y), D)
    tgt = s.target_matrix(st.T, st.Y, T1, Y1)
    losses = s.train(X, torch.as_tensor(tgt).cuda(), epochs=epochs, batch=2048, lr=1e-3, log=log)
    return s, losses


def test_surrogate_step_53sp(torch_gpu):
    """Surrogate-accelerated step on the 53-species workload: trained on the
    GPU integrator's own results, trajectories accepted only through the FP64
    implicit-residual and error test, the rest integrated by the plain BDF.
    Against the network-free proposal (every node at the initial state) under
    the same test, the network's trajectories must start the test's Newton
    iterations with smaller corrections.  Acceptance is decided by the error
    test, which at rtol 1e-6 only near-frozen reactors pass: the integrator's
    own states at the nodes, the proposal no network can improve on, must be
    accepted within a point of "hold", and the network's acceptance too.
    Accepted end states must agree with the plain BDF's within the
    integrator's own scale of differences, and rejected reactors are the
    plain BDF's."""
    from zerork_hip import capi, mechgen, surrogate
    torch = torch_gpu
    m = capi.Mech(*mech_paths("synth_53sp"))
    print("case surrogate 53sp", flush=True)
    s, losses = _train_surrogate(torch, capi, surrogate, mechgen, m, 131072, 11, epochs=12)
    assert losses[-1] < losses[0]
    n = 8192
    st = mechgen.workload_states("53sp", mechgen.workload_mechanism("53sp"), n, seed=12)
    rng = np.random.default_rng(12)
    dt = np.exp(rng.uniform(np.log(1e-6), np.log(1e-4), n))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    opts = capi.default_options(rel_tol=1e-6, abs_tol=1e-12)
    Tr, Pr, Yr = t(st.T), t(st.P), t(st.Y)
    _select(torch, surrogate, m, opts, Tr, Pr, Yr, t(dt))
    Tr, Yr = Tr.cpu().numpy(), Yr.cpu().numpy()
    out, conv, first = {}, {}, {}
    for prop in ("mlp", "hold"):
        T, P, Y = t(st.T), t(st.P), t(st.Y)
        acc, steps, res = s.solve(T, P, Y, t(dt), opts, proposal=prop)
        torch.cuda.synchronize()
        a = acc.cpu().numpy().astype(bool)
        res = res.cpu().numpy()
        conv[prop] = float(np.mean(res[:, 3] != 0))
        first[prop] = float(np.median(res[:, 1]))
        Tg, Yg = T.cpu().numpy(), Y.cpu().numpy()
        units = np.maximum(np.max(np.abs(Yg - Yr) / (1e-6 * np.abs(Yr) + 1e-12), axis=1),
                           np.abs(Tg - Tr) / (1e-6 * Tr))
        out[prop] = a.mean()
        print(f"{prop}: accepted {100 * a.mean():.1f} %; Newton failed "
              f"{100 * conv[prop]:.1f} %; median first correction {first[prop]:.3g}; accepted vs "
              f"plain BDF tol units p50/p99/max "
              f"{np.percentile(units[a], [50, 99, 100]).round(2) if a.any() else '-'}",
              flush=True)
        # rejected reactors went through the plain BDF: identical results
        np.testing.assert_array_equal(Yg[~a], Yr[~a])
        if a.any():
            assert np.percentile(units[a], 99) <= 100.0
    assert first["mlp"] < first["hold"], first
    # the perfect proposal: the plain integrator's states at the nodes
    Tn, Yn = [], []
    for f in surrogate.Surrogate.TAU:
        Tk, Pk, Yk = t(st.T), t(st.P), t(st.Y)
        _select(torch, surrogate, m, opts, Tk, Pk, Yk, t(dt * f))
        Tn.append(Tk)
        Yn.append(Yk)
    acc, res = s.accept(t(st.T), t(st.P), t(st.Y), torch.cat(Tn).contiguous(),
                        torch.cat(Yn).contiguous(), t(dt), opts)
    torch.cuda.synchronize()
    out["bdf nodes"] = float(acc.cpu().numpy().mean())
    print(f"bdf nodes: accepted {100 * out['bdf nodes']:.1f} %", flush=True)
    assert abs(out["bdf nodes"] - out["hold"]) <= 0.01, out
    assert abs(out["mlp"] - out["hold"]) <= 0.01, out
