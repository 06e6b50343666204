#!/usr/bin/env python
"""Benchmark of the batched reactor-chemistry path on one MI355X.

Workload (one "step" = one CFD step of the plugin path): 65,536
constant-pressure reactors of the seeded synthetic 53-species / 325-reaction
mechanism (C, H, O, N, Ar; 25 % third-body, 15 % Troe fall-off), T bimodal
(50 % U[700,950] K, 50 % U[1050,1400] K), P logU[1,50] atm, Y Dirichlet(0.5)
over 8 majors, integrated over dt = 1e-4 s with the FP64 BDF at rtol 1e-6,
atol 1e-12.  Inputs are resident in HBM; each step restores the initial
state from a device copy and integrates the whole batch.

Prints one JSON line (metric, value = reactors integrated per second, roofline
of the dominant kernel, the rate kernel's HBM bandwidth and a CPU baseline of
the oracle's CVODE restatement on a bounded sample of the same workload).
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
    p.add_argument("--reactors", type=int, default=65536)
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def pmc_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if any."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(mech, therm, st, dt, rtol, atol, seconds):
    """Oracle CVODE restatement on the host cores over a bounded sample."""
    import numpy as np
    import oracle
    o = oracle.Oracle(mech, therm)
    threads = max(1, min(16, os.cpu_count() or 1))
    opts = oracle.default_opts(rtol=rtol, atol=atol)
    # probe cost on a small sample, then size the timed sample
    n0 = min(len(st.T), 2 * threads)
    t0 = time.perf_counter()
    o.integrate(st.T[:n0], st.P[:n0], st.Y[:n0], dt, opts, nthreads=threads)
    per = (time.perf_counter() - t0) / n0
    n = int(max(n0, min(len(st.T), seconds / max(per, 1e-9))))
    # an evenly spaced sample keeps both temperature modes represented
    idx = np.linspace(0, len(st.T) - 1, n).astype(int)
    t0 = time.perf_counter()
    _, _, stats = o.integrate(st.T[idx], st.P[idx], st.Y[idx], dt, opts, nthreads=threads)
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "reactors/s", "cores": threads, "kind": "port",
            "sample": f"{n} evenly spaced reactors of the same 65,536-reactor batch, "
                      f"{el:.1f} s on {threads} threads (OpenMP, one reactor per thread)",
            "failed": int(np.sum(stats["nsteps"] < 0))}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from zerork_hip import capi, mechgen, workcount

    dev = torch.device("cuda:0")
    mech = os.path.join(ROOT, "tests", "golden", "mech", "synth_53sp.mech")
    therm = os.path.join(ROOT, "tests", "golden", "mech", "synth_53sp.therm")
    m = capi.Mech(mech, therm)
    mg = mechgen.workload_mechanism("53sp")
    n = a.reactors
    st = mechgen.workload_states("53sp", mg, n, seed=2)
    dt, rtol, atol = 1.0e-4, 1.0e-6, 1.0e-12
    nsp = m.nsp
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    T0, P, Y0 = t(st.T), t(st.P), t(st.Y)
    T, Y = T0.clone(), Y0.clone()
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    stats = torch.zeros(n * 8, dtype=torch.int32, device=dev)
    opts = capi.default_options(rel_tol=rtol, abs_tol=atol)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def one_step(with_stats=False):
        T.copy_(T0)
        Y.copy_(Y0)
        m.integrate(n, dt, T, P, Y, nsp, opts, steps, stats if with_stats else None, stream=sp)

    log(f"workload ready: {n} reactors, {nsp} species; warmup {a.warmup}")
    for _ in range(a.warmup):
        one_step()
    torch.cuda.synchronize()
    log("warmup done; timing")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * a.steps + 2)]
    ev[0].record(stream)
    for k in range(a.steps):
        T.copy_(T0)
        Y.copy_(Y0)
        ev[2 * k + 1].record(stream)
        m.integrate(n, dt, T, P, Y, nsp, opts, steps, None, stream=sp)
        ev[2 * k + 2].record(stream)
    torch.cuda.synchronize()
    total_ms = ev[0].elapsed_time(ev[2 * a.steps])
    kern_ms = sum(ev[2 * k + 1].elapsed_time(ev[2 * k + 2]) for k in range(a.steps)) / a.steps
    ms_per_step = total_ms / a.steps
    value = n / (ms_per_step / 1e3)
    nsteps = steps.cpu().numpy()
    failed = int(np.sum(nsteps < 0))

    log(f"timed {a.steps} steps: {ms_per_step:.2f} ms/step")
    # counters of the same integration (deterministic per reactor) for the
    # algorithmic FP64 operation count of the dominant kernel
    one_step(with_stats=True)
    torch.cuda.synchronize()
    s = stats.cpu().numpy().reshape(n, 8).astype(np.int64)
    W = workcount.Work(m)
    ops = W.bdf_ops(s[:, 1].sum(), s[:, 2].sum(), s[:, 3].sum(), s[:, 4].sum())
    achieved = ops / (kern_ms / 1e3) / 1e12
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", "bdf_kernel_53sp_65536_pmc.md"))
    roofline = {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                "traffic": traffic, "kernel": "bdf_kernel",
                "kernel_ms": round(kern_ms, 3), "fp64_ops_per_launch": int(ops)}

    # standalone rate kernel on the same batch: HBM bytes of one RHS evaluation
    ydot = torch.empty(n * (nsp + 1), dtype=torch.float64, device=dev)
    rop = torch.empty(n * m.nstep, dtype=torch.float64, device=dev)
    for _ in range(2):
        m.rates(n, T0, P, Y0, nsp, 1.0e22, ydot, rop, stream=sp)
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        m.rates(n, T0, P, Y0, nsp, 1.0e22, ydot, rop, stream=sp)
    e1.record(stream)
    torch.cuda.synchronize()
    rate_ms = e0.elapsed_time(e1) / reps
    rbytes = workcount.rate_bytes_per_reactor(nsp, m.nstep) * n
    rate_gbps = rbytes / (rate_ms / 1e3) / 1e9
    rate = {"kernel": "rates_kernel", "bound": "hbm", "achieved": round(rate_gbps, 1),
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(rate_gbps / HBM_PEAK_GBPS, 4),
            "bytes_per_launch": int(rbytes), "kernel_ms": round(rate_ms, 4),
            "fp64_ops_per_launch": int(W.rhs_ops() * n)}

    log(f"rate kernel {rate_ms:.3f} ms, {rate_gbps:.0f} GB/s")
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(mech, therm, st, dt, rtol, atol, a.cpu_seconds)

    if cpu:
        log(f"cpu baseline {cpu['value']:.1f} reactors/s on {cpu['cores']} threads")
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "reactors/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded mechanism and states)",
        "config": {"workload": "plain FP64 BDF, 53 species / 325 reactions x 65,536 "
                               "constant-pressure reactors, dt 1e-4 s, rtol 1e-6, atol 1e-12",
                   "reactors": n, "species": nsp, "steps_per_reaction_set": m.nstep,
                   "method": "plain_fp64", "failed_reactors": failed,
                   "mean_bdf_steps": round(float(nsteps[nsteps > 0].mean()), 1)},
        "roofline": roofline,
        "rate_kernel": rate,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
