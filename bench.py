#!/usr/bin/env python
"""Benchmark of the batched reactor-chemistry path on one MI355X.

Workload (one "step" = one CFD step of the plugin path), by default: 65,536
constant-pressure reactors of the seeded synthetic 53-species / 325-reaction
mechanism (C, H, O, N, Ar; 25 % third-body, 15 % Troe fall-off), T bimodal
(50 % U[700,950] K, 50 % U[1050,1400] K), P logU[1,50] atm, Y Dirichlet(0.5)
over 8 majors, integrated over dt = 1e-4 s with the BDF at rtol 1e-6,
atol 1e-12.  Inputs are resident in HBM; each step restores the initial
state from a device copy and integrates the whole batch.

Each timed step replays a HIP graph holding the outer step (restore the
batch, reset the work counter, integrate); the integrator kernel is also timed
alone with HIP events on its stream for the roofline.
Both methods are timed: the plain FP64 integrator and the chosen one (FP32
Newton matrix and inverse; rates, residual, error test and state in FP64).
Prints one JSON line: value = the chosen method's reactors per second, the
plain method beside it, their end-state agreement, the roofline of the
integrator kernel, the rate kernel's HBM bandwidth and a CPU baseline of the
oracle's CVODE restatement on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "zero-rk_amd"))

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = ("reactors integrated/sec at rtol 1e-6: chosen vs plain FP64 GPU; "
          "rate-kernel HBM GB/s")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", choices=["53sp", "10sp", "160sp", "2000sp"], default="53sp",
                   help="53sp (default, the headline workload); 10sp, 160sp and 2000sp "
                        "(sparse path) are the other seeded workloads, reported the same way")
    p.add_argument("--reactors", type=int, default=0,
                   help="batch size (default: the workload's own, 65,536 / 1,024 / 16,384)")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--copy-bytes", type=int, default=0,
                   help="only run a device-to-device copy of this many bytes (the "
                        "calibration kernel for the FETCH_SIZE/WRITE_SIZE counters)")
    return p.parse_args()


def copy_calibration(nbytes):
    """One plain device copy of a known size, for scaling the HBM counters."""
    import torch
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8}), flush=True)


def pmc_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if any."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mech, therm, st, dt, rtol, atol, seconds):
    """Oracle CVODE restatement on the host cores over a bounded sample: all
    host threads (up to the box's 16-core share), then one thread."""
    import numpy as np
    import oracle
    o = oracle.Oracle(mech, therm)
    threads = max(1, min(16, os.cpu_count() or 1))
    opts = oracle.default_opts(rtol=rtol, atol=atol)
    # probe cost on a small sample, then size the timed samples
    n0 = min(len(st.T), 2 * threads)
    t0 = time.perf_counter()
    o.integrate(st.T[:n0], st.P[:n0], st.Y[:n0], dt, opts, nthreads=threads)
    per = (time.perf_counter() - t0) / n0
    n = int(max(n0, min(len(st.T), seconds / max(per, 1e-9))))
    # an evenly spaced sample keeps both temperature modes represented
    idx = np.linspace(0, len(st.T) - 1, n).astype(int)
    t0 = time.perf_counter()
    To, Yo, stats = o.integrate(st.T[idx], st.P[idx], st.Y[idx], dt, opts, nthreads=threads)
    el = time.perf_counter() - t0
    # one thread: a smaller evenly spaced sample (about a third of the time)
    n1 = int(max(4, min(len(st.T), n * (seconds / 3.0) / (el * threads))))
    idx1 = np.linspace(0, len(st.T) - 1, n1).astype(int)
    t0 = time.perf_counter()
    o.integrate(st.T[idx1], st.P[idx1], st.Y[idx1], dt, opts, nthreads=1)
    el1 = time.perf_counter() - t0
    return {"value": n / el, "unit": "reactors/s", "cores": threads, "kind": "port",
            "_idx": idx, "_end": (To, Yo),
            "sample": f"{n} evenly spaced reactors of the same {len(st.T):,}-reactor batch, "
                      f"{el:.1f} s on {threads} threads (OpenMP, one reactor per thread)",
            "one_core": {"value": n1 / el1, "unit": "reactors/s",
                         "sample": f"{n1} evenly spaced reactors, {el1:.1f} s on 1 thread"},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "failed": int(np.sum(stats["nsteps"] < 0))}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.copy_bytes:
        copy_calibration(a.copy_bytes)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from zerork_hip import capi, mechgen, workcount

    dev = torch.device("cuda:0")
    wl = a.workload
    mech = os.path.join(ROOT, "tests", "golden", "mech", f"synth_{wl}.mech")
    therm = os.path.join(ROOT, "tests", "golden", "mech", f"synth_{wl}.therm")
    m = capi.Mech(mech, therm)
    mg = mechgen.workload_mechanism(wl)
    n = a.reactors or mechgen.WORKLOAD_NR[wl]
    st = mechgen.workload_states(wl, mg, n, seed=2)
    dt, rtol, atol = mechgen.WORKLOAD_DT[wl], 1.0e-6, 1.0e-12
    nsp = m.nsp
    sparse = nsp > 191     # above 191 species: the sparse path (FP32 LU only)
    has_plain = not sparse
    sp_nnz = sp_lev = sp_flops = 0
    if sparse:
        sp_nnz, sp_lev, sp_flops = m.sparse_plan_info()
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    T0, P0, Y0 = t(st.T), t(st.P), t(st.Y)
    T, P, Y = T0.clone(), P0.clone(), Y0.clone()
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    stats = torch.zeros(n * 8, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    W = workcount.Work(m)

    def run(fp32, Tstart=None):
        """warmup + K timed steps of one method; returns its measurements.
        Tstart: only integrate once from these temperatures (no timing)."""
        opts = capi.default_options(rel_tol=rtol, abs_tol=atol, fp32_preconditioner=fp32)

        def restore():
            T.copy_(T0)
            P.copy_(P0)
            Y.copy_(Y0)

        if Tstart is not None:
            restore()
            T.copy_(Tstart)
            m.integrate(n, dt, T, P, Y, nsp, opts, steps, None, stream=sp)
            torch.cuda.synchronize()
            return {"T": T.cpu().numpy(), "Y": Y.cpu().numpy()}

        for _ in range(max(1, a.warmup)):
            restore()
            m.integrate(n, dt, T, P, Y, nsp, opts, steps, None, stream=sp)
        torch.cuda.synchronize()
        log("warmup done")
        # the outer step (restore + integrate) captured once in a HIP graph
        graph = torch.cuda.CUDAGraph()
        cs = torch.cuda.Stream()
        cs.wait_stream(stream)
        with torch.cuda.stream(cs):
            with torch.cuda.graph(graph, stream=cs):
                restore()
                m.integrate(n, dt, T, P, Y, nsp, opts, steps, None,
                            stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for k in range(a.steps):
            graph.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        total_ms = e0.elapsed_time(e1)
        log(f"timed steps done: {total_ms / a.steps:.1f} ms/step")
        # the integrator kernel alone, eager, with HIP events on its stream
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        restore()
        ev[0].record(stream)
        m.integrate(n, dt, T, P, Y, nsp, opts, steps, None, stream=sp)
        ev[1].record(stream)
        torch.cuda.synchronize()
        kern_ms = ev[0].elapsed_time(ev[1])
        nst = steps.cpu().numpy()
        Tend, Yend = T.cpu().numpy(), Y.cpu().numpy()
        # counters of the same integration (deterministic per reactor) for the
        # algorithmic FP64 operation count of the integrator kernel
        restore()
        m.integrate(n, dt, T, P, Y, nsp, opts, steps, stats, stream=sp)
        torch.cuda.synchronize()
        s = stats.cpu().numpy().reshape(n, 8).astype(np.int64)
        if sparse:
            # FP64 RHS and Jacobian assembly, FP32 sparse factorisations and solves
            nfe, nje, nset, nni = (int(s[:, k].sum()) for k in (1, 2, 3, 4))
            ops = (nfe + 2 * nje) * W.rhs_ops() + nset * 2 * sp_flops + nni * 2 * sp_nnz
        else:
            ops = W.bdf_ops(s[:, 1].sum(), s[:, 2].sum(), s[:, 3].sum(), s[:, 4].sum())
        return {"ms_per_step": total_ms / a.steps, "kern_ms": kern_ms, "ops": ops,
                "failed": int(np.sum(nst < 0)), "mean_steps": float(nst[nst > 0].mean()),
                "T": Tend, "Y": Yend, "waves": m.resident_waves(opts)}

    log(f"workload ready: {n} reactors, {nsp} species; warmup {a.warmup}, steps {a.steps}")
    plain = run(0) if has_plain else None
    if plain:
        log(f"plain FP64: {plain['ms_per_step']:.2f} ms/step")
    chosen = run(1)
    log(f"FP32 preconditioner: {chosen['ms_per_step']:.2f} ms/step")

    def tol_units(A, B):
        """per-reactor max difference of two end states in tolerance units of B"""
        return np.maximum(np.max(np.abs(A["Y"] - B["Y"]) / (rtol * np.abs(B["Y"]) + atol), axis=1),
                          np.abs(A["T"] - B["T"]) / (rtol * B["T"]))

    def quantiles(u):
        return {"p50": round(float(np.percentile(u, 50)), 3),
                "p99": round(float(np.percentile(u, 99)), 3), "max": round(float(u.max()), 3)}

    agreement = sensitivity = None
    if plain:
        # end-state agreement of the two methods, in tolerance units of the plain
        # one, beside the plain method's own spread under a one-ulp change of T
        agreement = quantiles(tol_units(chosen, plain))
        sensitivity = quantiles(tol_units(run(0, Tstart=torch.nextafter(
            T0, torch.full_like(T0, float("inf")))), plain))
    value = n / (chosen["ms_per_step"] / 1e3)
    achieved = chosen["ops"] / (chosen["kern_ms"] / 1e3) / 1e12
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", f"bdf_kernel_{wl}_{n}_pmc.md"))
    roofline = {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                "traffic": traffic,
                "kernel": ("bdf_kernel<float, false>" if nsp <= 63 else
                           "bdf_sparse_kernel" if sparse else "bdf_team_kernel"),
                "kernel_ms": round(chosen["kern_ms"], 3),
                "fp64_ops_per_launch": int(chosen["ops"]),
                "resident_waves": chosen["waves"]}

    if sparse:
        roofline["sparse_lu"] = {"entries": sp_nnz, "levels": sp_lev,
                                 "multiply_adds_per_factorisation": sp_flops,
                                 "note": "ops = FP64 RHS and Jacobian + FP32 LU and solves"}
    rate = cpu = None
    rate = rate_kernel(m, n, nsp, T0, P0, Y0, sp, stream, W)
    vs_oracle = None
    if not a.no_cpu and not sparse:
        cpu = cpu_baseline(mech, therm, st, dt, rtol, atol, a.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.1f} reactors/s on {cpu['cores']} threads")
        # end states of both methods against the FP64 oracle on the same sample
        idx = cpu.pop("_idx")
        To, Yo = cpu.pop("_end")

        def errors(R):
            rel = np.abs(R["Y"][idx] - Yo) / np.maximum(np.abs(Yo), atol / rtol)
            relT = np.abs(R["T"][idx] - To) / To
            u = np.maximum(np.max(np.abs(R["Y"][idx] - Yo) / (rtol * np.abs(Yo) + atol), axis=1),
                           relT / rtol)
            return {"max_rel_err_Y": float(rel.max()), "rms_rel_err_Y": float(np.sqrt(np.mean(rel**2))),
                    "max_rel_err_T": float(relT.max()), "rms_rel_err_T": float(np.sqrt(np.mean(relT**2))),
                    "tol_units_p50_p99_max": [round(float(np.percentile(u, q)), 3) for q in (50, 99, 100)]}

        vs_oracle = {"reactors": int(len(idx)),
                     "note": "relative errors with |Y| floored at atol/rtol; tolerance units "
                             "|y - y_oracle| / (rtol |y_oracle| + atol)",
                     "chosen": errors(chosen)}
        if plain:
            vs_oracle["plain_fp64"] = errors(plain)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "reactors/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(chosen["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded mechanism and states)",
        "config": {"workload": f"{nsp} species / {m.nrxn} reactions x {n:,} constant-pressure "
                               f"reactors, dt {dt:g} s, rtol 1e-6, atol 1e-12; chosen method = "
                               + ("sparse FP32 LU of the Newton matrix (FP64 rates, Jacobian, "
                                  "residual, error test, state)" if sparse else
                                  "FP32 preconditioner (FP64 rates, residual, error test, "
                                  "state)"),
                   "reactors": n, "species": nsp, "steps_per_reaction_set": m.nstep,
                   "method": "sparse_fp32_lu" if sparse else "fp32_preconditioner",
                   "failed_reactors": chosen["failed"],
                   "mean_bdf_steps": round(chosen["mean_steps"], 1)},
        "plain_fp64": None if not plain else {
            "value": round(n / (plain["ms_per_step"] / 1e3), 1),
            "ms_per_step": round(plain["ms_per_step"], 3),
            "kernel_ms": round(plain["kern_ms"], 3),
            "fp64_ops_per_launch": int(plain["ops"]),
            "tflops": round(plain["ops"] / (plain["kern_ms"] / 1e3) / 1e12, 4),
            "failed_reactors": plain["failed"],
            "mean_bdf_steps": round(plain["mean_steps"], 1)},
        "chosen_vs_plain": round(plain["ms_per_step"] / chosen["ms_per_step"], 3) if plain else None,
        "chosen_vs_plain_end_state_tol_units": agreement,
        "plain_fp64_one_ulp_T_sensitivity_tol_units": sensitivity,
        "roofline": roofline,
        "rate_kernel": rate,
        "end_state_error_vs_fp64_oracle": vs_oracle,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def rate_kernel(m, n, nsp, T0, P0, Y0, sp, stream, W):
    """standalone rate kernel on the same batch: HBM bytes of one RHS evaluation"""
    import torch
    from zerork_hip import workcount
    dev = T0.device
    ydot = torch.empty(n * (nsp + 1), dtype=torch.float64, device=dev)
    rop = torch.empty(n * m.nstep, dtype=torch.float64, device=dev)
    for _ in range(2):
        m.rates(n, T0, P0, Y0, nsp, 1.0e22, ydot, rop, stream=sp)
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        m.rates(n, T0, P0, Y0, nsp, 1.0e22, ydot, rop, stream=sp)
    e1.record(stream)
    torch.cuda.synchronize()
    rate_ms = e0.elapsed_time(e1) / reps
    rbytes = workcount.rate_bytes_per_reactor(nsp, m.nstep) * n
    rate_gbps = rbytes / (rate_ms / 1e3) / 1e9
    # one wavefront per reactor while its rate arrays fit in a quarter of LDS,
    # else one workgroup per reactor (the kernels' own rule)
    block = 8 * 4 * (3 * nsp + 2 + m.nstep + 1) > 160 * 1024
    rate = {"kernel": "rates_block_kernel<false>" if block else "rates_kernel<false>",
            "bound": "hbm", "achieved": round(rate_gbps, 1),
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(rate_gbps / HBM_PEAK_GBPS, 4),
            "bytes_per_launch": int(rbytes), "kernel_ms": round(rate_ms, 4),
            "fp64_ops_per_launch": int(W.rhs_ops() * n)}
    log(f"rate kernel {rate_ms:.3f} ms, {rate_gbps:.0f} GB/s")
    return rate


if __name__ == "__main__":
    main()
