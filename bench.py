#!/usr/bin/env python3
"""Benchmark: full FP64 pf-FRG flow on one MI355X.

Workload (default): SU(2) simple cubic J1 = 1, J2 = 0.3, lattice range 6
(L = 23 representative sites, 377 sites in range), 64 log-spaced positive
frequencies in [0.002, 100] (N_f = 133,120 frequency triples), cutoff 100 ->
0.005 with step 0.985 (656 cutoffs, 655 derivative evaluations).  Synthetic in
the sense that it is generated from the task parameters; the initial vertex is
the model's couplings, exactly as the reference initialises a flow.

One step = one iteration of the reference's flow loop (src/SpinParser.cpp:141-160):
computeStep (dSigma + dGamma), correlation measurement, divergence check,
finalizeStep (in-place Euler).  The timed steps walk the cutoff schedule from the
start; K = 655 is the full flow and yields the full-schedule average of
dGamma/dLambda evaluations per second.  Warm-up steps run first on the same
context, which is then reset to the initial condition.

Roofline: the dGamma kernel is timed with HIP events on the context's stream;
achieved = algorithmic FLOP of the timed evaluations (exact integrand-call
counts x per-call work model) / summed kernel time, against the MI355X FP64
peak.  The CPU baseline times the reference-arithmetic (float) CPU oracle with
OpenMP over items on a bounded sample of the same workload and extrapolates by
integrand-call counts over the schedule.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from spinparser_amd.capi import KERNEL_DGAMMA, FlowContext  # noqa: E402
from spinparser_amd.problem import su2_problem  # noqa: E402
from spinparser_amd.task import make_task_xml, parse_task  # noqa: E402
from spinparser_amd.workload import compulsory_bytes, evaluation_calls, evaluation_flops  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "cubic-j1j2-r6-nw64": dict(lattice="cubic", lattice_range=6, model="cubic-j1j2heisenberg",
                               couplings={"j1": 1.0, "j2": 0.3}, freq_min=0.002, freq_max=100.0, freq_count=64,
                               cutoff_max=100.0, cutoff_min=0.005, cutoff_step=0.985),
    "square-r4-nw32": dict(lattice="square", lattice_range=4, model="square-heisenberg", couplings={"j": 1.0},
                           freq_min=0.005, freq_max=50.0, freq_count=32, cutoff_max=50.0, cutoff_min=0.01,
                           cutoff_step=0.98),
}


def build_problem(name):
    xml = make_task_xml(symmetry="SU2", **WORKLOADS[name])
    return su2_problem(parse_task(xml, is_text=True))


def cpu_baseline(prob, target_s=15.0):
    """Reference-arithmetic CPU oracle on a bounded item sample at five cutoffs.

    Items are taken as 8 contiguous chunks at random offsets per cutoff; the
    measured time per algorithmic FLOP (from the oracle's exact per-channel
    integrand-call counts) is extrapolated to the whole schedule."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    o = Oracle(prob)
    v2, v4s, v4d = prob.initial_state(np.float32)
    dv2 = np.zeros(prob.nw, dtype=np.float32)
    ncut = len(prob.cutoffs)
    ks = [int(round(x)) for x in np.linspace(0, ncut - 2, 5)]
    rng = np.random.default_rng(1)
    chunks = 8
    frac = 1.0 / 1024
    for rep in range(2):
        t_tot, f_tot, items_tot = 0.0, 0.0, 0
        for k in ks:
            n_chunk = max(threads, int(prob.n_freq * frac / chunks))
            for _ in range(chunks):
                start = int(rng.integers(0, prob.n_freq - n_chunk))
                t0 = time.perf_counter()
                _, _, calls = o.items(float(prob.cutoffs[k]), v2, dv2, v4s, v4d, start, start + n_chunk,
                                      dtype=np.float32, nthreads=threads)
                t_tot += time.perf_counter() - t0
                f_tot += prob.flops_per_evaluation(*calls.tolist()) - prob.flops_per_evaluation(0, 0, 0)
                items_tot += n_chunk
        if rep == 0:
            frac = min(1.0, frac * max(1.0, target_s / max(t_tot, 1e-3)))
    sec_per_flop = t_tot / f_tot
    flow_flops = sum(evaluation_flops(prob, prob.cutoffs[k]) for k in range(ncut - 1))
    flow_s = sec_per_flop * flow_flops
    return {
        "value": (ncut - 1) / flow_s, "unit": "dGamma/dLambda evals/s", "cores": threads, "kind": "port",
        "sample": "float32 CPU oracle (reference arithmetic, OpenMP guided over items) on %d items at each of "
                  "%d cutoffs (%.1f s measured, %.3g GFLOP/s); full-flow time %.0f s extrapolated by the "
                  "algorithmic FLOP of all %d evaluations" % (items_tot // len(ks), len(ks), t_tot,
                                                                 f_tot / t_tot / 1e9, flow_s, ncut - 1),
    }


def run_step(ctx, prob, k):
    ctx.compute_step()
    ctx.measure_correlation()
    if ctx.is_diverged():
        raise RuntimeError("flow diverged at cutoff index %d" % k)
    ctx.finalize_step(float(prob.cutoffs[k + 1]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=655)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cubic-j1j2-r6-nw64", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs one flow per GPU and is measured on one GPU only; "
              "--gpus %d is not supported (the multi-GPU sharding is a design, not a build)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    prob = build_problem(args.workload)
    nsteps_flow = len(prob.cutoffs) - 1
    ctx = FlowContext(prob)
    init = prob.initial_state(np.float64)

    ctx.set_state(float(prob.cutoffs[0]), *init)
    for w in range(args.warmup):
        run_step(ctx, prob, w % nsteps_flow)
    ctx.set_state(float(prob.cutoffs[0]), *init)
    ctx.synchronize()

    ctx.profile(True)
    flops = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        k = i % nsteps_flow
        if k == 0 and i > 0:
            ctx.set_state(float(prob.cutoffs[0]), *init)
        run_step(ctx, prob, k)
        flops += evaluation_flops(prob, prob.cutoffs[k])
        if (i + 1) % 100 == 0:
            print("# step %d / %d  cutoff %.5g  %.1f s" % (i + 1, args.steps, prob.cutoffs[k],
                                                          time.perf_counter() - t0), file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    dg_ms, dg_n = ctx.profile_get(KERNEL_DGAMMA)
    ctx.profile(False)
    ctx.close()

    achieved = flops / (dg_ms * 1e-3) / 1e12
    hbm_gbs = compulsory_bytes(prob) * args.steps / elapsed / 1e9
    out = {
        "metric": "dGamma/dLambda evals/s (full-schedule average, FP64 flow)",
        "value": args.steps / elapsed,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (task-file generated flow, initial vertex = model couplings)",
        "config": {"workload": args.workload, "L": prob.L, "nw": prob.nw, "n_freq_triples": prob.n_freq,
                   "overlap_terms": prob.overlap_total, "cutoffs": len(prob.cutoffs),
                   "step": "computeStep + correlation + isDiverged + finalizeStep"},
        "hbm_gbs_compulsory": hbm_gbs,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_dgamma", "launches": dg_n, "avg_launch_ms": dg_ms / max(dg_n, 1),
                     "algorithmic_tflop_total": flops / 1e12},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
