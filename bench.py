#!/usr/bin/env python3
"""Benchmark of the pf-FRG flow engine on one MI355X.

Mode "flow" (default).  Workload: SU(2) simple cubic J1 = 1, J2 = 0.3, lattice
range 6 (L = 23 representative sites, 377 sites in range), 64 log-spaced
positive frequencies in [0.002, 100] (N_f = 133,120 frequency triples), cutoff
100 -> 0.005 with step 0.985 (656 cutoffs, 655 derivative evaluations).
Synthetic in the sense that it is generated from the task parameters; the
initial vertex is the model's couplings, exactly as the reference initialises a
flow.

One step = one iteration of the reference's flow loop (src/SpinParser.cpp:141-160):
computeStep (dSigma + dGamma), correlation measurement, divergence check,
finalizeStep (in-place Euler).  The timed steps walk the cutoff schedule from the
start; K = 655 is the full flow and yields the full-schedule average of
dGamma/dLambda evaluations per second (``value``, FP64 form).  Beside it:

* ``mixed``: the same flow in the mixed-precision form (FP32 frequency
  integrals with compensated sums; FP64 state, Euler, self-energy and
  correlations), timed the same way, with the max-relative error against the
  FP64 flow of every output: correlations (every record), final state, and the
  derivative on a common state;
* ``fixed_cutoff``: evals/s at the first, middle and last cutoff with the
  algorithmic work F(cutoff) of one evaluation;
* ``roofline``: the dGamma kernel timed with HIP events on the context's
  stream; achieved = algorithmic FLOP of the timed evaluations (exact
  integrand-call counts x per-call work model) / summed kernel time, against
  the MI355X FP64 peak; ``traffic`` = HBM bytes per launch from rocprofv3
  FETCH_SIZE / WRITE_SIZE (table below);
* ``cpu_baseline``: the reference-arithmetic CPU oracle with OpenMP over items
  on a bounded sample of the same workload, extrapolated by integrand-call
  counts over the schedule.

* ``tri``: the TRI core on kagome J = 1, D = 0.2 (range 5, 48 frequencies,
  cutoff 50 -> 0.01 x 0.98): evaluations at five cutoffs over the schedule and
  the full-schedule rate integrated by algorithmic FLOP, with its own CPU
  baseline (float TRI oracle on a bounded item sample);
* ``xyz``: the same for the XYZ core on the honeycomb Heisenberg-Kitaev model
  J = 0.1, K = -1 (range 8: L = 23, 1,525 overlap terms; 64 frequencies in
  [0.005, 50], cutoff 50 -> 0.01 x 0.98).

Modes "tri-flow" / "xyz-flow": the whole TRI / XYZ flow (421 evaluations)
measured end to end, with roofline and CPU baseline.

Mode "sweep": the served J2/J1 phase-diagram sweep (64 SU(2) cubic flows, J2 ~
U[0,1], range ~ U{3..7}, N_w ~ U{32,48,64}, Poisson arrivals, mixed precision,
admission by device footprint) against the same flows run one at a time in
FP64; flows/hour, p50/p99 latency and the correlation error of every served flow
against its FP64 run.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from spinparser_amd.capi import (KERNEL_DGAMMA, SP_PRECISION_FP64, SP_PRECISION_MIXED,  # noqa: E402
                                 FlowContext, TriFlowContext, XyzFlowContext, footprint_bytes)
from spinparser_amd.problem import su2_problem, tri_problem, xyz_problem  # noqa: E402
from spinparser_amd.task import make_task_xml, parse_task  # noqa: E402
from spinparser_amd.workload import compulsory_bytes, evaluation_flops  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
HBM_BYTES = 288 * 1024 ** 3

WORKLOADS = {
    "cubic-j1j2-r6-nw64": dict(lattice="cubic", lattice_range=6, model="cubic-j1j2heisenberg",
                               couplings={"j1": 1.0, "j2": 0.3}, freq_min=0.002, freq_max=100.0, freq_count=64,
                               cutoff_max=100.0, cutoff_min=0.005, cutoff_step=0.985),
    "square-r4-nw32": dict(lattice="square", lattice_range=4, model="square-heisenberg", couplings={"j": 1.0},
                           freq_min=0.005, freq_max=50.0, freq_count=32, cutoff_max=50.0, cutoff_min=0.01,
                           cutoff_step=0.98),
}

# TRI core: kagome J = 1 with out-of-plane Dzyaloshinskii-Moriya D = 0.2, range
# 5 (L = 19, 771 overlap terms), 48 log-spaced frequencies in [0.005, 50],
# cutoff 50 -> 0.01 with step 0.98 (422 cutoffs)
TRI_WORKLOAD = dict(lattice="kagome", lattice_range=5, model="kagome-DM", couplings={"j": 1.0, "d": 0.2},
                    freq_min=0.005, freq_max=50.0, freq_count=48, cutoff_max=50.0, cutoff_min=0.01, cutoff_step=0.98)
# XYZ core: honeycomb Heisenberg-Kitaev J = 0.1, K = -1, range 8 (L = 23, 1,525
# overlap terms, all six spin permutations among the overlap transformations),
# 64 log-spaced frequencies in [0.005, 50], cutoff 50 -> 0.01 with step 0.98
XYZ_WORKLOAD = dict(lattice="honeycomb", lattice_range=8, model="honeycomb-kitaev", couplings={"j": 0.1, "k": -1.0},
                    freq_min=0.005, freq_max=50.0, freq_count=64, cutoff_max=50.0, cutoff_min=0.01, cutoff_step=0.98)

# the cores measured beside the SU(2) flow: workload, problem builder, context,
# oracle class name and dGamma kernel name
CORES = {
    "tri": dict(label="TRI", workload="kagome-dm-r5-nw48", spec=TRI_WORKLOAD, symmetry="TRI", build=tri_problem,
                context=TriFlowContext, oracle="TriOracle", kernel="k_tri_dgamma"),
    "xyz": dict(label="XYZ", workload="honeycomb-kitaev-r8-nw64", spec=XYZ_WORKLOAD, symmetry="XYZ",
                build=xyz_problem, context=XyzFlowContext, oracle="XyzOracle", kernel="k_xyz_dgamma"),
}

# HBM bytes per k_dgamma launch over the full flow, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (KiB per dispatch, averaged over the 655 dispatches
# of the full flow).  Each counter is multiplied by the ratio of known bytes to
# counter reading for a plain streaming kernel of known size (k_euler reads
# Gamma and dGamma, writes Gamma: FETCH_SIZE reads half the known bytes, scale
# 2.00; WRITE_SIZE reads them exactly, scale 1.00).  The profile summaries
# under profiles/ hold the runs and the commands.
TRAFFIC_KIB = {
    ("cubic-j1j2-r6-nw64", "f64"): {"fetch_kib": 738093.0, "write_kib": 47871.7,
                                    "fetch_scale": 2.00, "write_scale": 1.00},
    ("square-r4-nw32", "f64"): {"fetch_kib": 6045.33, "write_kib": 2376.01,
                                "fetch_scale": 2.00, "write_scale": 1.00},
    ("kagome-dm-r5-nw48", "f64"): {"fetch_kib": 17817400.0, "write_kib": 134064.0,
                                   "fetch_scale": 2.00, "write_scale": 1.00},
    ("honeycomb-kitaev-r8-nw64", "f64"): {"fetch_kib": 3350080.0, "write_kib": 96774.9,
                                          "fetch_scale": 2.00, "write_scale": 1.00},
}


def host_threads():
    """CPU threads of the baselines: OMP_NUM_THREADS when set, else the CPUs
    this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def build_problem(name):
    xml = make_task_xml(symmetry="SU2", **WORKLOADS[name])
    return su2_problem(parse_task(xml, is_text=True))


def build_core_problem(core):
    c = CORES[core]
    xml = make_task_xml(symmetry=c["symmetry"], **c["spec"])
    return c["build"](parse_task(xml, is_text=True))


def max_rel(a, b):
    """Normwise relative error max|a-b| / max|b|."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = float(np.abs(b).max())
    num = float(np.abs(a - b).max())
    return num / den if den > 0 else num


def cpu_baseline(prob, target_s=15.0):
    """Reference-arithmetic CPU oracle on a bounded item sample at five cutoffs.

    Items are taken as 8 contiguous chunks at random offsets per cutoff; the
    measured time per algorithmic FLOP (from the oracle's exact per-channel
    integrand-call counts) is extrapolated to the whole schedule."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    threads = host_threads()
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
        "value": (ncut - 1) / flow_s, "unit": "dGamma/dLambda evals/s", "cores": threads,
        "cpu_model": cpu_model(), "kind": "port",
        "sample": "float32 CPU oracle (reference arithmetic, OpenMP guided over items) on %d items at each of "
                  "%d cutoffs (%.1f s measured, %.3g GFLOP/s); full-flow time %.0f s extrapolated by the "
                  "algorithmic FLOP of all %d evaluations" % (items_tot // len(ks), len(ks), t_tot,
                                                                 f_tot / t_tot / 1e9, flow_s, ncut - 1),
    }


def run_step(ctx, prob, k, records=None):
    ctx.compute_step()
    zz, dd = ctx.measure_correlation()
    if records is not None:
        records.append((zz, dd))
    if ctx.is_diverged():
        raise RuntimeError("flow diverged at cutoff index %d" % k)
    ctx.finalize_step(float(prob.cutoffs[k + 1]))


def flow_leg(prob, precision, steps, warmup, label, queued=True):
    """Time `steps` iterations of the flow loop from the first cutoff.  queued:
    the loop runs through sp_run_flow (computeStep, takeMeasurements, isDiverged
    and finalizeStep per step, queued without a host round trip each; records
    read back every step); otherwise one ABI call per stage and step."""
    nsteps_flow = len(prob.cutoffs) - 1
    ctx = FlowContext(prob, precision)
    init = prob.initial_state(np.float64)
    ctx.set_state(float(prob.cutoffs[0]), *init)
    for w in range(warmup):
        run_step(ctx, prob, w % nsteps_flow)
    ctx.set_state(float(prob.cutoffs[0]), *init)
    # algorithmic FLOP per evaluation, counted before the timed region
    step_flops = [evaluation_flops(prob, prob.cutoffs[k]) for k in range(nsteps_flow)]
    ctx.synchronize()
    ctx.profile(True)
    flops = 0.0
    records = []
    t0 = time.perf_counter()
    if queued:
        done = 0
        while done < steps:
            n = min(steps - done, nsteps_flow)
            if done > 0:
                ctx.set_state(float(prob.cutoffs[0]), *init)
            recs, diverged = ctx.run_flow(prob.cutoffs[:n + 1])
            if diverged:
                raise RuntimeError("flow diverged at cutoff index %d" % (len(recs) - 1))
            records = recs
            flops += sum(step_flops[:n])
            done += n
            print("# %s step %d / %d  %.1f s" % (label, done, steps, time.perf_counter() - t0), file=sys.stderr,
                  flush=True)
    else:
        for i in range(steps):
            k = i % nsteps_flow
            if k == 0 and i > 0:
                ctx.set_state(float(prob.cutoffs[0]), *init)
                records = []
            run_step(ctx, prob, k, records)
            flops += step_flops[k]
            if (i + 1) % 100 == 0:
                print("# %s step %d / %d  cutoff %.5g  %.1f s" % (label, i + 1, steps, prob.cutoffs[k],
                                                                 time.perf_counter() - t0), file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    dg_ms, dg_n = ctx.profile_get(KERNEL_DGAMMA)
    ctx.profile(False)
    final = ctx.get_state() if steps % nsteps_flow == 0 else None
    ctx.close()
    return dict(elapsed=elapsed, dg_ms=dg_ms, dg_n=dg_n, flops=flops, records=records, final=final)


def fixed_cutoff(prob, precision, reps=5):
    """evals/s at the first, middle and last evaluated cutoff.  The work of an
    evaluation depends on the cutoff only (integration windows), so the state
    is the initial condition placed at that cutoff."""
    ncut = len(prob.cutoffs)
    init = prob.initial_state(np.float64)
    out = []
    ctx = FlowContext(prob, precision)
    for k in (0, (ncut - 2) // 2, ncut - 2):
        ctx.set_state(float(prob.cutoffs[k]), *init)
        ctx.compute_step()
        ctx.synchronize()
        ctx.profile(True)
        ms0, n0 = ctx.profile_get(KERNEL_DGAMMA)  # the totals accumulate over the context's life
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.compute_step()
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / reps
        ms1, n1 = ctx.profile_get(KERNEL_DGAMMA)
        ctx.profile(False)
        f = evaluation_flops(prob, prob.cutoffs[k])
        kern = (ms1 - ms0) * 1e-3 / max(n1 - n0, 1)
        out.append({"cutoff": float(prob.cutoffs[k]), "gflop_per_eval": f / 1e9, "evals_per_s": 1.0 / wall,
                    "k_dgamma_ms": kern * 1e3, "tflops_k_dgamma": f / kern / 1e12})
    ctx.close()
    return out


def derivative_errors(prob, state):
    """Mixed vs FP64 derivative on one common state."""
    flows = []
    for prec in (SP_PRECISION_FP64, SP_PRECISION_MIXED):
        ctx = FlowContext(prob, prec)
        ctx.set_state(*state)
        ctx.compute_step()
        flows.append(ctx.get_flow())
        ctx.close()
    names = ("dSigma", "dGamma_spin", "dGamma_density")
    return {n: max_rel(a, b) for n, a, b in zip(names, flows[1], flows[0])}


def traffic_entry(workload, dtype):
    t = TRAFFIC_KIB.get((workload, dtype))
    if t is None:
        return None
    return (t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024.0


def core_leg(core, samples=5, reps=2):
    """TRI or XYZ core on its workload: dGamma evaluations at `samples` cutoffs
    spread over the schedule (state = initial condition placed at the cutoff;
    the work of an evaluation depends on the cutoff only), and the
    full-schedule rate obtained by integrating the measured time per
    algorithmic FLOP over the algorithmic work of every evaluation."""
    c = CORES[core]
    p = build_core_problem(core)
    ncut = len(p.cutoffs)
    ks = [int(round(x)) for x in np.linspace(0, ncut - 2, samples)]
    v2, v4 = p.initial_state(np.float64)
    ctx = c["context"](p)
    kkey = c["kernel"] + "_ms"
    tkey = "tflops_" + c["kernel"]
    pts = []
    for k in ks:
        ctx.set_state(float(p.cutoffs[k]), v2, v4)
        ctx.compute_step()
        ctx.synchronize()
        ctx.profile(True)
        ms0, n0 = ctx.profile_get(KERNEL_DGAMMA)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.compute_step()
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / reps
        ms1, n1 = ctx.profile_get(KERNEL_DGAMMA)
        ctx.profile(False)
        f = evaluation_flops(p, p.cutoffs[k])
        kern = (ms1 - ms0) * 1e-3 / max(1, n1 - n0)
        pts.append({"cutoff": float(p.cutoffs[k]), "index": k, "gflop_per_eval": f / 1e9, "evals_per_s": 1.0 / wall,
                    kkey: kern * 1e3, tkey: f / kern / 1e12})
    ctx.close()
    # seconds per FLOP, piecewise linear in the cutoff index between samples
    idx = np.array([q["index"] for q in pts], dtype=float)
    spf = np.array([1.0 / (q["evals_per_s"] * q["gflop_per_eval"] * 1e9) for q in pts])
    flops = np.array([evaluation_flops(p, p.cutoffs[k]) for k in range(ncut - 1)])
    est = float(np.sum(flops * np.interp(np.arange(ncut - 1), idx, spf)))
    return {"workload": c["workload"], "L": p.L, "nw": p.nw, "overlap_terms": p.overlap_total,
            "n_freq_triples": p.n_freq, "cutoffs": ncut, "samples": pts,
            "full_flow_algorithmic_tflop": float(flops.sum()) / 1e12,
            "est_full_flow_s": est, "est_evals_per_s": (ncut - 1) / est,
            "method": "%d evaluations at each of %d cutoffs, integrated over the schedule by algorithmic FLOP"
                      % (reps, samples)}


def core_cpu_baseline(core, target_s=10.0):
    """float CPU oracle of the TRI / XYZ core (reference arithmetic, OpenMP
    guided over items) on a bounded item sample at five cutoffs, extrapolated
    by algorithmic FLOP."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    c = CORES[core]
    threads = host_threads()
    p = build_core_problem(core)
    o = getattr(oracle, c["oracle"])(p)
    v2, v4 = p.initial_state(np.float32)
    dv2 = np.zeros(p.nw, dtype=np.float32)
    ncut = len(p.cutoffs)
    ks = [int(round(x)) for x in np.linspace(0, ncut - 2, 5)]
    rng = np.random.default_rng(2)
    n_items = max(threads, 64)
    for rep in range(2):
        t_tot, f_tot = 0.0, 0.0
        for k in ks:
            start = int(rng.integers(0, p.n_freq - n_items))
            t0 = time.perf_counter()
            _, calls = o.items(float(p.cutoffs[k]), v2, dv2, v4, start, start + n_items, dtype=np.float32,
                               nthreads=threads)
            t_tot += time.perf_counter() - t0
            f_tot += p.flops_per_evaluation(*calls.tolist()) - p.flops_per_evaluation(0, 0, 0)
        if rep == 0:
            n_items = int(min(p.n_freq // 2, n_items * max(1.0, target_s / max(t_tot, 1e-3))))
    flow_flops = sum(evaluation_flops(p, p.cutoffs[k]) for k in range(ncut - 1))
    flow_s = t_tot / f_tot * flow_flops
    return {"value": (ncut - 1) / flow_s, "unit": "dGamma/dLambda evals/s", "cores": threads,
        "cpu_model": cpu_model(), "kind": "port",
            "sample": "float32 %s CPU oracle on %d items at each of 5 cutoffs (%.1f s measured, %.3g GFLOP/s); "
                      "full flow %.0f s extrapolated by algorithmic FLOP" % (c["label"], n_items, t_tot,
                                                                          f_tot / t_tot / 1e9, flow_s)}


def mode_core_flow(args, core):
    """The whole TRI / XYZ flow on the GPU, one step = computeStep + correlation
    + isDiverged + finalizeStep."""
    c = CORES[core]
    p = build_core_problem(core)
    nsteps_flow = len(p.cutoffs) - 1
    steps = min(args.steps, nsteps_flow)
    ctx = c["context"](p)
    v2, v4 = p.initial_state(np.float64)
    ctx.set_state(float(p.cutoffs[0]), v2, v4)
    for w in range(args.warmup):
        ctx.compute_step()
        ctx.measure_correlation()
        ctx.finalize_step(float(p.cutoffs[w + 1]))
    ctx.set_state(float(p.cutoffs[0]), v2, v4)
    flops = sum(evaluation_flops(p, p.cutoffs[k]) for k in range(steps))  # counted before the timed region
    ctx.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.compute_step()
        ctx.measure_correlation()
        if ctx.is_diverged():
            raise RuntimeError("%s flow diverged at cutoff index %d" % (c["label"], k))
        ctx.finalize_step(float(p.cutoffs[k + 1]))
        if (k + 1) % 20 == 0:
            print("# %s step %d / %d  %.1f s" % (core, k + 1, steps, time.perf_counter() - t0), file=sys.stderr,
                  flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    dg_ms, dg_n = ctx.profile_get(KERNEL_DGAMMA)
    ctx.close()
    achieved = flops / (dg_ms * 1e-3) / 1e12
    out = {"metric": "%s dGamma/dLambda evals/s (full-schedule average, FP64 flow)" % c["label"],
           "value": steps / elapsed,
           "unit": "evals/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": elapsed * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (task-file generated flow, initial vertex = model couplings)",
           "config": {"workload": c["workload"], "L": p.L, "nw": p.nw, "overlap_terms": p.overlap_total,
                      "cutoffs": len(p.cutoffs), "step": "computeStep + correlation + isDiverged + finalizeStep"},
           "roofline": {"bound": "valu_fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_entry(c["workload"], "f64"),
                        "kernel": c["kernel"],
                        "launches": dg_n, "avg_launch_ms": dg_ms / max(dg_n, 1),
                        "algorithmic_tflop_total": flops / 1e12},
           "cpu_baseline": None}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = core_cpu_baseline(core, args.cpu_seconds)
    print(json.dumps(out))


def sample_indices(nsteps_flow, steps):
    """`steps` cutoff indices stratified over the schedule: the midpoints of
    `steps` equal strata of the evaluation indices 0 .. nsteps_flow - 1."""
    return [min(nsteps_flow - 1, int((i + 0.5) * nsteps_flow / steps)) for i in range(steps)]


def full_schedule_seconds(step_flops, ks, t_steps):
    """Whole-schedule time from timed samples: the seconds per algorithmic FLOP
    measured at the sampled evaluation indices `ks`, interpolated linearly in
    the index (held constant beyond the first and last sample), times the FLOP
    of every evaluation of the schedule."""
    step_flops = np.asarray(step_flops, dtype=np.float64)
    spf = np.array([t / step_flops[k] for t, k in zip(t_steps, ks)])
    return float(np.sum(step_flops * np.interp(np.arange(len(step_flops)), ks, spf)))


def stratified_leg(prob, precision, steps, warmup, label):
    """Exactly `steps` iterations of the flow loop (computeStep, correlation,
    isDiverged, finalizeStep; one ABI call per stage) at cutoffs stratified
    over the whole schedule, each on the state the flow itself reaches there.

    An untimed pass of the flow (sp_run_flow) stops at every sampled cutoff
    and copies its state into a context of its own, so the timed region holds
    the K evaluations only.  The time of an evaluation depends on the cutoff
    (its integration windows), not on the vertex values; the full-schedule time
    is the algorithmic FLOP of every evaluation times the measured seconds per
    FLOP, interpolated linearly in the cutoff index between the samples."""
    nsteps_flow = len(prob.cutoffs) - 1
    ks = sample_indices(nsteps_flow, steps)
    walker = FlowContext(prob, precision)
    walker.set_state(float(prob.cutoffs[0]), *prob.initial_state(np.float64))
    ctxs, pos = [], 0
    for k in ks:
        if k > pos:
            _, div = walker.run_flow(prob.cutoffs[pos:k + 1])
            if div:
                raise RuntimeError("%s: flow diverged before cutoff index %d" % (label, k))
            pos = k
        c = FlowContext(prob, precision)
        c.set_state(*walker.get_state())
        ctxs.append(c)
    for _ in range(warmup):  # untimed: the walker sits at the last sampled cutoff
        walker.compute_step()
        walker.measure_correlation()
        walker.is_diverged()
    walker.close()
    step_flops = [evaluation_flops(prob, prob.cutoffs[k]) for k in range(nsteps_flow)]
    for c in ctxs:
        c.synchronize()
        c.profile(True)
    t_steps = []
    t0 = time.perf_counter()
    for c, k in zip(ctxs, ks):
        ts = time.perf_counter()
        c.compute_step()
        c.measure_correlation()
        if c.is_diverged():
            raise RuntimeError("%s: flow diverged at cutoff index %d" % (label, k))
        c.finalize_step(float(prob.cutoffs[k + 1]))
        c.synchronize()
        t_steps.append(time.perf_counter() - ts)
    elapsed = time.perf_counter() - t0
    dg_ms, dg_n = 0.0, 0
    for c in ctxs:
        ms, n = c.profile_get(KERNEL_DGAMMA)
        dg_ms += ms
        dg_n += n
        c.close()
    print("# %s: %d stratified steps %.3f s" % (label, steps, elapsed), file=sys.stderr, flush=True)
    est = full_schedule_seconds(step_flops, ks, t_steps)
    return dict(elapsed=elapsed, dg_ms=dg_ms, dg_n=dg_n, flops=float(sum(step_flops[k] for k in ks)),
                est_full_s=est, indices=ks)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def records_max_rel(ra, rb):
    """Per-channel max over records of the normwise relative error."""
    n = min(len(ra), len(rb))
    return [max(max_rel(a[c], b[c]) for a, b in zip(ra[:n], rb[:n])) for c in range(2)]


def square_leg(args):
    """The CPU-runnable configuration (square lattice J = 1, range 4, 32
    frequencies in [0.005, 50], cutoff 50 -> 0.01 x 0.98, 421 evaluations) end
    to end: the FP64 GPU flow and the float CPU oracle flow (the reference's
    arithmetic, OpenMP guided over items), each the whole cutoff loop with its
    correlation records."""
    prob = build_problem("square-r4-nw32")
    nsteps = len(prob.cutoffs) - 1
    r = flow_leg(prob, SP_PRECISION_FP64, nsteps, 2, "square fp64")
    out = {"workload": "square-r4-nw32", "L": prob.L, "nw": prob.nw, "n_freq_triples": prob.n_freq,
           "cutoffs": len(prob.cutoffs), "gpu_evals_per_s": nsteps / r["elapsed"], "gpu_flow_s": r["elapsed"],
           "gpu_k_dgamma_avg_ms": r["dg_ms"] / max(r["dg_n"], 1),
           "gpu_tflops_k_dgamma": r["flops"] / (r["dg_ms"] * 1e-3) / 1e12}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        threads = host_threads()
        t0 = time.perf_counter()
        zz, dd, n, _ = Oracle(prob).run_flow(np.float32, nthreads=threads)
        cpu_s = time.perf_counter() - t0
        cpu_recs = [(zz[k].astype(np.float64), dd[k].astype(np.float64)) for k in range(min(n, nsteps))]
        ezz, edd = records_max_rel(r["records"], cpu_recs)
        out["cpu_baseline"] = {"value": nsteps / cpu_s, "unit": "dGamma/dLambda evals/s", "cores": threads,
                               "cpu_model": cpu_model(), "kind": "port", "flow_s": cpu_s,
                               "sample": "the whole float32 CPU oracle flow (reference arithmetic, OpenMP guided "
                                         "over items), %d evaluations and correlation records" % nsteps}
        out["gpu_vs_cpu"] = out["gpu_evals_per_s"] / out["cpu_baseline"]["value"]
        out["max_rel_err_corr_gpu_fp64_vs_cpu_float"] = {"zz": ezz, "dd": edd}
    return out


def mode_flow(args):
    prob = build_problem(args.workload)
    nsteps_flow = len(prob.cutoffs) - 1
    stratified = args.steps < nsteps_flow
    if stratified:
        r64 = stratified_leg(prob, SP_PRECISION_FP64, args.steps, args.warmup, "fp64")
        value = nsteps_flow / r64["est_full_s"]
        how = ("%d evaluations stratified over the %d-evaluation schedule (each on the flow's own state at that "
               "cutoff), full-schedule time = algorithmic FLOP of every evaluation x measured s/FLOP interpolated "
               "over the cutoff index" % (args.steps, nsteps_flow))
    else:
        r64 = flow_leg(prob, SP_PRECISION_FP64, args.steps, args.warmup, "fp64")
        value = args.steps / r64["elapsed"]
        how = "whole flows from the first cutoff, %d evaluations timed" % args.steps
    elapsed = r64["elapsed"]
    # the whole FP64 flow through sp_run_flow, beside the headline; the roofline
    # describes the dGamma kernel over this whole flow (the launches a profile of
    # the full-flow command averages over)
    full = r64 if not stratified else flow_leg(prob, SP_PRECISION_FP64, nsteps_flow, 0, "fp64 full flow")
    achieved = full["flops"] / (full["dg_ms"] * 1e-3) / 1e12
    traffic = traffic_entry(args.workload, "f64")
    avg_ms = full["dg_ms"] / max(full["dg_n"], 1)
    out = {
        "metric": "dGamma/dLambda evals/s (full-schedule average, FP64 flow)",
        "value": value,
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
                   "step": "computeStep + correlation + isDiverged + finalizeStep",
                   "timing": how},
        "hbm_gbs_compulsory": compulsory_bytes(prob) / (avg_ms * 1e-3) / 1e9,
        "roofline": {"bound": "valu_fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "k_dgamma", "launches": full["dg_n"], "avg_launch_ms": avg_ms,
                     "algorithmic_tflop_total": full["flops"] / 1e12,
                     "peak_note": "MI355X FP64 vector FMA peak (equal to the FP64 matrix peak); the kernel issues "
                                  "no MFMA: gathers and sparse site sums",
                     "window": "the whole flow (%d evaluations, HIP events on the context stream); traffic = "
                               "mean HBM bytes per launch of the same flow (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                               "scaled)" % nsteps_flow},
        "cpu_baseline": None,
    }
    if traffic is not None:
        out["roofline"]["traffic_gbs"] = traffic / (avg_ms * 1e-3) / 1e9
    out["full_flow"] = {"evals_per_s": nsteps_flow / full["elapsed"] if stratified else value,
                        "flow_s": full["elapsed"], "k_dgamma_avg_ms": avg_ms,
                        "tflops_k_dgamma": achieved, "frac": achieved / FP64_PEAK_TFLOPS,
                        "loop": "sp_run_flow (steps queued, records read back every step, divergence checked "
                                "on the device)"}
    if stratified:
        w_ms = r64["dg_ms"] / max(r64["dg_n"], 1)
        out["timed_window"] = {"k_dgamma_avg_ms": w_ms, "launches": r64["dg_n"],
                               "tflops_k_dgamma": r64["flops"] / (r64["dg_ms"] * 1e-3) / 1e12,
                               "cutoff_indices": r64["indices"]}
    if not args.no_step_api and not stratified:
        # the same flow with one ABI call per stage and step (the FrgCore call pattern)
        rs = flow_leg(prob, SP_PRECISION_FP64, args.steps, 0, "fp64 step api", queued=False)
        out["step_api"] = {"evals_per_s": args.steps / rs["elapsed"], "ms_per_step": rs["elapsed"] * 1e3 / args.steps,
                           "records_identical_to_queued": bool(
                               len(rs["records"]) == len(r64["records"]) and
                               all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                                   for a, b in zip(rs["records"], r64["records"])))}
    if not args.no_fixed:
        out["fixed_cutoff"] = fixed_cutoff(prob, SP_PRECISION_FP64)
    if not args.no_mixed:
        if stratified:
            m = stratified_leg(prob, SP_PRECISION_MIXED, args.steps, 0, "mixed")
            mrate = nsteps_flow / m["est_full_s"]
        else:
            m = flow_leg(prob, SP_PRECISION_MIXED, args.steps, 0, "mixed")
            mrate = args.steps / m["elapsed"]
        r32 = m if not stratified else flow_leg(prob, SP_PRECISION_MIXED, nsteps_flow, 0, "mixed full flow")
        mixed = {"evals_per_s": mrate, "k_dgamma_avg_ms": m["dg_ms"] / max(m["dg_n"], 1),
                 "speedup_vs_fp64": mrate / value,
                 "full_flow_evals_per_s": nsteps_flow / r32["elapsed"] if stratified else mrate}
        if full["records"] and len(full["records"]) == len(r32["records"]):
            mixed["max_rel_err_corr_zz"], mixed["max_rel_err_corr_dd"] = records_max_rel(r32["records"],
                                                                                       full["records"])
            mixed["records_compared"] = len(full["records"])
        if full["final"] is not None and r32["final"] is not None:
            names = ("Sigma", "Gamma_spin", "Gamma_density")
            mixed["max_rel_err_final_state"] = {n: max_rel(a, b) for n, a, b in
                                                zip(names, r32["final"][1:], full["final"][1:])}
            mixed["max_rel_err_derivative_final_state"] = derivative_errors(prob, full["final"])
        init = prob.initial_state(np.float64)
        mid = float(prob.cutoffs[(nsteps_flow - 1) // 2])
        mixed["max_rel_err_derivative_initial_state_mid_cutoff"] = derivative_errors(prob, (mid, *init))
        out["mixed"] = mixed
    legs = [c for c, off in (("tri", args.no_tri), ("xyz", args.no_xyz)) if not off]
    for c in legs:
        out[c] = core_leg(c)
    if not args.no_square:
        out["square"] = square_leg(args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob, args.cpu_seconds)
        for c in legs:
            out[c]["cpu_baseline"] = core_cpu_baseline(c, min(10.0, args.cpu_seconds))
    if not args.no_sweep:
        out["sweep"] = sweep_leg(args)
    print(json.dumps(out))


def sweep_leg(args):
    """The served J2/J1 sweep against the same flows run one at a time in FP64
    (sp_run_flow).  Returns throughput, latency, and the correlation error of
    every served flow against its FP64 run."""
    from spinparser_amd.serve import PrecisionSchedule, SweepServer, run_alone, sweep_requests
    t0 = time.perf_counter()
    reqs = sweep_requests(args.flows, seed=args.seed, rate_per_s=args.rate)
    print("# %d sweep requests built in %.1f s" % (len(reqs), time.perf_counter() - t0), file=sys.stderr, flush=True)
    precision = SP_PRECISION_FP64 if args.sweep_fp64 else SP_PRECISION_MIXED
    schedule = None if (args.sweep_fp64 or args.fp64_below <= 0) else PrecisionSchedule(args.fp64_below)
    fps = [footprint_bytes(r.problem, precision) for r in reqs]
    budget = min(HBM_BYTES - (16 << 30), int(args.budget_gb * (1 << 30)))
    srv = SweepServer(budget_bytes=budget, max_flows=args.flows, precision=precision, schedule=schedule)
    t0 = time.perf_counter()
    served = srv.run(reqs, log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    t_served = time.perf_counter() - t0
    srv.close()
    lat = np.array([r.latency for r in served])
    form = ("FP64" if precision == SP_PRECISION_FP64 else
            "mixed" if schedule is None else
            "mixed above %.3g J, FP64 below" % schedule.fp64_below)
    out = {"workload": "cubic-j1j2-sweep-%d" % args.flows, "seed": args.seed, "rate_per_s": args.rate,
           "precision": form, "flows": len(reqs),
           "flows_per_hour": len(served) / t_served * 3600.0, "sweep_wall_s": t_served,
           "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
           "evaluations": int(sum(r.steps for r in served)),
           "fp64_fraction_of_evaluations": srv.fp64_steps / max(1, srv.fp64_steps + srv.mixed_steps),
           "budget_bytes": budget, "footprint_total_bytes": int(sum(fps)), "footprint_max_bytes": int(max(fps)),
           "peak_flows_resident": srv.peak_flows, "peak_bytes_resident": srv.peak_bytes,
           "mean_batch": float(np.mean(srv.batch_sizes)), "diverged": int(sum(r.diverged for r in served))}
    if args.no_alone:
        return out
    t_alone = 0.0
    worst = {"zz": 0.0, "dd": 0.0, "flow": None}
    worst_div = {"zz": 0.0, "dd": 0.0}
    same_end = 0
    per_flow = []
    for r, sv in zip(reqs, served):
        a = run_alone(r, SP_PRECISION_FP64, queued=True)
        t_alone += a.latency
        same = len(a.zz) == len(sv.zz) and a.diverged == sv.diverged
        same_end += int(same)
        n = min(len(a.zz), len(sv.zz))
        # a flow that diverges has its final records dominated by the blow-up,
        # where relative comparisons carry no information: its records up to
        # ten before the divergence are reported separately
        n_cmp = max(0, n - 10) if (a.diverged or sv.diverged) else n
        ezz = max((max_rel(x, y) for x, y in zip(sv.zz[:n_cmp], a.zz[:n_cmp])), default=0.0)
        edd = max((max_rel(x, y) for x, y in zip(sv.dd[:n_cmp], a.dd[:n_cmp])), default=0.0)
        if a.diverged or sv.diverged:
            worst_div["zz"], worst_div["dd"] = max(worst_div["zz"], ezz), max(worst_div["dd"], edd)
        elif max(ezz, edd) > max(worst["zz"], worst["dd"]):
            worst = {"zz": max(worst["zz"], ezz), "dd": max(worst["dd"], edd), "flow": r.ident}
        else:
            worst["zz"], worst["dd"] = max(worst["zz"], ezz), max(worst["dd"], edd)
        per_flow.append({"flow": r.ident, "j2": round(r.params["j2"], 4), "range": r.params["range"],
                         "nw": r.params["nw"], "records": len(sv.zz), "records_fp64": len(a.zz),
                         "diverged": bool(sv.diverged), "max_rel_err": max(ezz, edd)})
        print("# FP64 alone flow %d %s: %.2f s, %d records, diverged %s (served: %d, %s), err %.2e" % (
            r.ident, r.params, a.latency, len(a.zz), a.diverged, len(sv.zz), sv.diverged, max(ezz, edd)),
            file=sys.stderr, flush=True)
    out["fp64_one_at_a_time"] = {"wall_s": t_alone, "flows_per_hour": len(reqs) / t_alone * 3600.0,
                                 "loop": "sp_run_flow, one flow after another"}
    out["speedup_vs_fp64_one_at_a_time"] = out["flows_per_hour"] / out["fp64_one_at_a_time"]["flows_per_hour"]
    if precision == SP_PRECISION_MIXED:
        out["max_rel_err_corr_vs_fp64"] = {"zz": worst["zz"], "dd": worst["dd"], "worst_flow": worst["flow"],
                                           "flows": len(reqs) - int(sum(r.diverged for r in served))}
        out["max_rel_err_corr_vs_fp64_diverging_flows_before_last_10"] = worst_div
    out["flows_same_end_as_fp64"] = same_end
    out["per_flow"] = per_flow
    return out


def mode_sweep(args):
    sw = sweep_leg(args)
    out = {
        "metric": "served sweep flows/hour (continuous batching, %s)" % sw["precision"],
        "value": sw["flows_per_hour"],
        "unit": "flows/hour",
        "n_gpus": 1,
        "steps": sw["evaluations"],
        "warmup": 0,
        "ms_per_step": sw["sweep_wall_s"] * 1e3 / max(1, sw["evaluations"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.sweep_fp64 else "f32+f64",
        "data": "synthetic (cubic J1-J2 flows drawn from PCG64 seed %d)" % args.seed,
        "config": {"workload": sw["workload"], "seed": args.seed, "rate_per_s": args.rate,
                   "ranges": [3, 4, 5, 6, 7], "nws": [32, 48, 64]},
        "sweep": sw,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=655)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="flow", choices=("flow", "sweep", "tri-flow", "xyz-flow"))
    ap.add_argument("--workload", default="cubic-j1j2-r6-nw64", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-fixed", action="store_true")
    ap.add_argument("--no-step-api", action="store_true")
    ap.add_argument("--no-tri", action="store_true")
    ap.add_argument("--no-xyz", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--flows", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--rate", type=float, default=1.0, help="Poisson arrival rate of the sweep (flows/s)")
    ap.add_argument("--budget-gb", type=float, default=256.0)
    ap.add_argument("--no-alone", action="store_true")
    ap.add_argument("--no-square", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-fp64", action="store_true", help="serve the sweep in the FP64 form")
    ap.add_argument("--fp64-below", type=float, default=3.0,
                    help="mixed sweep: FP64 form at cutoffs below this multiple of the coupling scale (0: never)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU and is measured on one GPU only; "
              "--gpus %d is not supported (the multi-GPU sharding is a design, not a build)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "flow":
        mode_flow(args)
    elif args.mode == "tri-flow":
        mode_core_flow(args, "tri")
    elif args.mode == "xyz-flow":
        mode_core_flow(args, "xyz")
    else:
        mode_sweep(args)


if __name__ == "__main__":
    main()
