"""Served sweep: continuous batching of many resident flows on one GPU.

A request is one flow (its problem tables, built host-side from a task file).
The loop keeps every admitted flow's vertex state resident in HBM and advances
all of them together, one cutoff per iteration, with one launch per kernel for
the whole batch (``sp_group_step``).  Requests are admitted in arrival order as
soon as they have arrived and their device footprint fits the memory budget;
a flow leaves the batch when it reaches the end of its cutoff schedule (final
measurement) or diverges, which frees its slot and memory for the next request.

Each flow follows the reference's cutoff loop (``src/SpinParser.cpp:141-172``):
computeStep, correlation measurement, divergence check, finalizeStep, and a
final measurement at the last cutoff.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Tuple

import numpy as np

from .capi import SP_PRECISION_FP64, SP_PRECISION_MIXED, FlowContext, FlowGroup, footprint_bytes
from .problem import SU2Problem


def coupling_scale(prob: SU2Problem) -> float:
    """Largest coupling |J| of the flow (the initial vertex is J / normalization)."""
    return float(np.abs(prob.coupling_zz_f64).max()) * float(prob.normalization)


@dataclass
class PrecisionSchedule:
    """Precision form of a mixed-precision flow along its cutoff schedule: the
    FP64 form for cutoffs at or below ``fp64_below`` times the flow's coupling
    scale J, the mixed form above.  The state is FP64 in both forms, so the
    switch only changes how later derivatives are evaluated.  Measured on the
    J2/J1 sweep, rounding errors of the FP32 integrals made at cutoffs of the
    order of J and below are amplified by up to three orders of magnitude
    towards the end of the flow, where the susceptibility of flows close to an
    ordering instability grows; errors made at large cutoffs are not."""

    fp64_below: float = 3.0

    def precision(self, prob: SU2Problem, cutoff: float) -> int:
        return SP_PRECISION_FP64 if cutoff <= self.fp64_below * coupling_scale(prob) else SP_PRECISION_MIXED


@dataclass
class FlowRequest:
    ident: int
    problem: SU2Problem
    arrival: float                 # seconds after the start of the sweep
    params: dict = field(default_factory=dict)


@dataclass
class FlowResult:
    ident: int
    params: dict
    arrival: float
    admitted: float
    finished: float
    steps: int
    diverged: bool
    cutoffs: List[float]
    zz: List[np.ndarray]           # SU2CorZZ per measured cutoff
    dd: List[np.ndarray]           # SU2CorDD per measured cutoff

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class _Active:
    def __init__(self, req: FlowRequest, ctx: FlowContext, fp: int, admitted: float):
        self.req = req
        self.ctx = ctx
        self.prob = req.problem
        self.fp = fp
        self.k = 0
        self.admitted = admitted
        self.cutoffs: List[float] = []
        self.zz: List[np.ndarray] = []
        self.dd: List[np.ndarray] = []
        self.diverged = False


class SweepServer:
    """Request loop with footprint admission and continuous batching."""

    def __init__(self, budget_bytes: int, max_flows: int = 64, precision: int = SP_PRECISION_FP64,
                 clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
                 schedule: Optional[PrecisionSchedule] = None):
        self.budget = int(budget_bytes)
        self.max_flows = int(max_flows)
        self.precision = precision
        # flows created in the mixed form follow the schedule (None: mixed throughout)
        self.schedule = schedule if precision == SP_PRECISION_MIXED else None
        self.fp64_steps = 0
        self.mixed_steps = 0
        self.clock = clock
        self.sleep = sleep
        self.group = FlowGroup(self.max_flows)
        self.peak_flows = 0
        self.peak_bytes = 0
        self.batch_sizes: List[int] = []

    def close(self):
        self.group.close()

    def _admit(self, req: FlowRequest, now: float) -> _Active:
        fp = footprint_bytes(req.problem, self.precision)
        ctx = FlowContext(req.problem, self.precision)
        v2, v4s, v4d = req.problem.initial_state(np.float64)
        ctx.set_state(float(req.problem.cutoffs[0]), v2, v4s, v4d)
        return _Active(req, ctx, fp, now)

    def run(self, requests: List[FlowRequest], log: Optional[Callable[[str], None]] = None,
            log_every: int = 50) -> List[FlowResult]:
        pending = sorted(requests, key=lambda r: (r.arrival, r.ident))
        for r in pending:
            if footprint_bytes(r.problem, self.precision) > self.budget:
                raise ValueError("flow %d does not fit the memory budget on its own" % r.ident)
        active: List[_Active] = []
        done: List[FlowResult] = []
        used = 0
        t0 = self.clock()
        while pending or active:
            now = self.clock() - t0
            while (pending and pending[0].arrival <= now and len(active) < self.max_flows
                   and used + footprint_bytes(pending[0].problem, self.precision) <= self.budget):
                a = self._admit(pending.pop(0), now)
                used += a.fp
                active.append(a)
            self.peak_flows = max(self.peak_flows, len(active))
            self.peak_bytes = max(self.peak_bytes, used)
            if not active:
                self.sleep(max(0.0, pending[0].arrival - now))
                continue
            self.batch_sizes.append(len(active))
            if log and log_every > 0 and len(self.batch_sizes) % log_every == 0:
                log("iteration %d: %d active, %d pending, %d done, %.1f GB resident, t=%.1f s" % (
                    len(self.batch_sizes), len(active), len(pending), len(done), used / 1e9, now))
            for a in active:
                want = self.precision if self.schedule is None else \
                    self.schedule.precision(a.prob, float(a.prob.cutoffs[a.k]))
                if want != a.ctx.precision:
                    a.ctx.set_precision(want)
                    a.ctx.synchronize()
                if want == SP_PRECISION_FP64:
                    self.fp64_steps += 1
                else:
                    self.mixed_steps += 1
            new_cut = [float(a.prob.cutoffs[a.k + 1]) for a in active]
            outs, div = self.group.step([a.ctx for a in active], new_cut)
            finishing: List[_Active] = []
            for a, (zz, dd), d in zip(active, outs, div):
                a.cutoffs.append(float(a.prob.cutoffs[a.k]))
                a.zz.append(zz)
                a.dd.append(dd)
                if d:
                    a.diverged = True
                    finishing.append(a)
                    continue
                a.k += 1
                if a.k == len(a.prob.cutoffs) - 1:
                    finishing.append(a)
            last = [a for a in finishing if not a.diverged]
            if last:
                for a, (zz, dd) in zip(last, self.group.measure([a.ctx for a in last])):
                    a.cutoffs.append(float(a.prob.cutoffs[a.k]))
                    a.zz.append(zz)
                    a.dd.append(dd)
            if finishing:
                t_fin = self.clock() - t0
                for a in finishing:
                    a.ctx.close()
                    used -= a.fp
                    active.remove(a)
                    done.append(FlowResult(a.req.ident, a.req.params, a.req.arrival, a.admitted, t_fin, a.k,
                                           a.diverged, a.cutoffs, a.zz, a.dd))
                    if log:
                        log("flow %d done: %d steps, latency %.2f s, %d active" % (a.req.ident, a.k,
                                                                                t_fin - a.req.arrival,
                                                                                len(active)))
        return sorted(done, key=lambda r: r.ident)


def run_alone(req: FlowRequest, precision: int, schedule: Optional[PrecisionSchedule] = None,
              queued: bool = False) -> FlowResult:
    """The same flow driven through the single-flow entry points (no batching):
    one ABI call per stage and step, or (queued) the whole cutoff loop through
    sp_run_flow.  A mixed flow follows the precision schedule when one is given."""
    prob = req.problem
    ctx = FlowContext(prob, precision)
    v2, v4s, v4d = prob.initial_state(np.float64)
    ctx.set_state(float(prob.cutoffs[0]), v2, v4s, v4d)
    t0 = time.perf_counter()
    cutoffs, zzs, dds = [], [], []
    k = 0
    diverged = False
    if queued and schedule is None:
        recs, diverged = ctx.run_flow(prob.cutoffs)
        cutoffs = [float(c) for c in prob.cutoffs[:len(recs)]]
        zzs = [r[0] for r in recs]
        dds = [r[1] for r in recs]
        k = len(recs) - 1 if diverged else len(recs)
    while not diverged and k < len(prob.cutoffs) - 1:
        if schedule is not None and precision == SP_PRECISION_MIXED:
            want = schedule.precision(prob, float(prob.cutoffs[k]))
            if want != ctx.precision:
                ctx.set_precision(want)
        ctx.compute_step()
        zz, dd = ctx.measure_correlation()
        cutoffs.append(float(prob.cutoffs[k]))
        zzs.append(zz)
        dds.append(dd)
        if ctx.is_diverged():
            diverged = True
            break
        k += 1
        ctx.finalize_step(float(prob.cutoffs[k]))
    if not diverged:
        zz, dd = ctx.measure_correlation()
        cutoffs.append(float(prob.cutoffs[k]))
        zzs.append(zz)
        dds.append(dd)
    ctx.close()
    t1 = time.perf_counter()
    return FlowResult(req.ident, req.params, 0.0, 0.0, t1 - t0, k, diverged, cutoffs, zzs, dds)


def sweep_requests(n: int, seed: int, rate_per_s: float, ranges=(3, 4, 5, 6, 7), nws=(32, 48, 64),
                   fmin=0.002, fmax=100.0, cmax=100.0, cmin=0.005, cstep=0.985) -> List[FlowRequest]:
    """The J2/J1 phase-diagram sweep on the simple cubic lattice: J2 ~ U[0,1],
    range ~ U{ranges}, N_w ~ U{nws}, Poisson arrivals at the given rate (0: all
    at once), all drawn from one PCG64 seed.  The lattice tables depend on the
    range only (the J1-J2 couplings keep the full cubic symmetry for every J2),
    so the factory runs once per range and only the couplings are redone."""
    import dataclasses
    from .lattice import LatticeUnitCell, SpinModelUnitCell, model_couplings, new_lattice_model
    from .problem import su2_problem_from_lattice
    from .task import exponential_cutoffs, exponential_mesh
    rng = np.random.Generator(np.random.PCG64(seed))
    uc = LatticeUnitCell.from_resources("cubic")
    lattices = {}
    cutoffs = exponential_cutoffs(cmax, cmin, cstep)
    t = 0.0
    out = []
    for i in range(n):
        j2 = float(rng.uniform(0.0, 1.0))
        rr = int(rng.choice(ranges))
        nw = int(rng.choice(nws))
        t += float(rng.exponential(1.0 / rate_per_s)) if rate_per_s > 0 else 0.0
        if rr not in lattices:
            generic = SpinModelUnitCell.from_resources("cubic-j1j2heisenberg", {"j1": "1.0", "j2": "0.5"})
            lattices[rr] = new_lattice_model(uc, generic, rr)
        model = SpinModelUnitCell.from_resources("cubic-j1j2heisenberg", {"j1": "1.0", "j2": repr(j2)})
        lat = dataclasses.replace(lattices[rr], couplings=model_couplings(lattices[rr].sites[:lattices[rr].size],
                                                                           model))
        prob = su2_problem_from_lattice(exponential_mesh(fmin, fmax, nw), cutoffs, lat, {})
        out.append(FlowRequest(i, prob, t, {"j2": j2, "range": rr, "nw": nw}))
    return out
