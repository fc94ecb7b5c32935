"""Flow groups, mixed precision and the served sweep on the GPU.

* a group step equals the single-flow entry points on every flow of the group
  (same kernels, one launch for all flows; tolerance 1e-13 normwise);
* the mixed-precision form (FP32 frequency integrals with compensated sums,
  FP64 state / Euler / self-energy / measurement) stays within 1e-6 of the FP64
  flow on every correlation record, and its per-step derivative error is stated;
* the served sweep returns, for every flow, the same records as that flow run
  alone, while admitting flows by footprint under a memory budget.
"""
import sys

import numpy as np
import pytest

from helpers import cfg1_problem, max_rel, python_obs_problem, ref1_problem

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _records_close(a, b, tol):
    assert len(a.cutoffs) == len(b.cutoffs)
    worst = 0.0
    for za, zb, da, db in zip(a.zz, b.zz, a.dd, b.dd):
        worst = max(worst, max_rel(za, zb), max_rel(da, db))
    assert worst <= tol, worst
    return worst


def test_group_step_equals_single_flow():
    from spinparser_amd.capi import SP_PRECISION_FP64
    from spinparser_amd.serve import FlowRequest, SweepServer, run_alone
    probs = [cfg1_problem(nw=8), ref1_problem(), python_obs_problem()]
    reqs = [FlowRequest(i, p, 0.0) for i, p in enumerate(probs)]
    srv = SweepServer(budget_bytes=1 << 34, max_flows=4, precision=SP_PRECISION_FP64)
    mixed = srv.run(reqs)
    srv.close()
    assert srv.peak_flows == 3
    for r in reqs:
        alone = run_alone(r, SP_PRECISION_FP64)
        w = _records_close(mixed[r.ident], alone, 1e-13)
        _log("flow %d (L=%d nw=%d): group vs alone %.2g over %d records" % (r.ident, r.problem.L, r.problem.nw, w,
                                                                           len(alone.cutoffs)))


def test_group_step_two_lds_classes():
    """A group whose flows need different LDS plans issues one dGamma launch per
    class (cubic range 8, L = 41, N_w = 16: 41 KB even at one point per slot,
    above the quarter of the CU's LDS; square range 4: 28 KB); every flow still
    equals its single-flow run."""
    from spinparser_amd.capi import SP_PRECISION_FP64
    from spinparser_amd.serve import FlowRequest, SweepServer, run_alone
    from helpers import exp_problem
    big = exp_problem("cubic", 8, "cubic-j1j2heisenberg", {"j1": 1.0, "j2": 0.3}, 16, 0.002, 100.0, 100.0, 0.05, 0.9)
    assert big.L == 41
    probs = [cfg1_problem(nw=8), big]
    reqs = [FlowRequest(i, p, 0.0) for i, p in enumerate(probs)]
    srv = SweepServer(budget_bytes=1 << 34, max_flows=2, precision=SP_PRECISION_FP64)
    grouped = srv.run(reqs)
    srv.close()
    assert srv.peak_flows == 2
    for r in reqs:
        alone = run_alone(r, SP_PRECISION_FP64)
        w = _records_close(grouped[r.ident], alone, 1e-13)
        _log("flow %d (L=%d nw=%d): group vs alone %.2g over %d records" % (r.ident, r.problem.L, r.problem.nw, w,
                                                                           len(alone.cutoffs)))


def test_mixed_precision_error_vs_fp64():
    from spinparser_amd.capi import SP_PRECISION_FP64, SP_PRECISION_MIXED, FlowContext
    from spinparser_amd.serve import FlowRequest, run_alone
    from helpers import cfg2_problem
    prob = cfg2_problem(nw=16, lattice_range=4)
    req = FlowRequest(0, prob, 0.0)
    r64 = run_alone(req, SP_PRECISION_FP64)
    r32 = run_alone(req, SP_PRECISION_MIXED)
    worst_zz = max(max_rel(a, b) for a, b in zip(r32.zz, r64.zz))
    worst_dd = max(max_rel(a, b) for a, b in zip(r32.dd, r64.dd))
    _log("mixed vs FP64 correlations over %d cutoffs: zz %.3g dd %.3g" % (len(r64.cutoffs), worst_zz, worst_dd))
    assert worst_zz < 1e-6 and worst_dd < 1e-6
    # per-step derivative error on a common state
    c64 = FlowContext(prob, SP_PRECISION_FP64)
    c32 = FlowContext(prob, SP_PRECISION_MIXED)
    v2, v4s, v4d = prob.initial_state(np.float64)
    cut = float(prob.cutoffs[400])
    for c in (c64, c32):
        c.set_state(cut, v2, v4s, v4d)
        c.compute_step()
    f64, f32 = c64.get_flow(), c32.get_flow()
    e = [max_rel(a, b) for a, b in zip(f32, f64)]
    _log("mixed vs FP64 derivative at cutoff %.4g: dSigma %.3g dGs %.3g dGd %.3g" % (cut, *e))
    assert e[0] == 0.0 and e[1] < 1e-5 and e[2] < 1e-5
    c64.close()
    c32.close()


def test_served_sweep_alone_vs_mix_and_admission():
    from spinparser_amd.capi import SP_PRECISION_MIXED, footprint_bytes
    from spinparser_amd.serve import SweepServer, run_alone, sweep_requests
    reqs = sweep_requests(6, seed=11, rate_per_s=50.0, ranges=(2, 3), nws=(8, 12), cmax=100.0, cmin=0.05, cstep=0.9)
    fps = [footprint_bytes(r.problem, SP_PRECISION_MIXED) for r in reqs]
    budget = int(max(fps) + sorted(fps)[0])  # at most two flows resident
    srv = SweepServer(budget_bytes=budget, max_flows=8, precision=SP_PRECISION_MIXED)
    res = srv.run(reqs, log=_log)
    srv.close()
    assert srv.peak_bytes <= budget
    assert len(res) == len(reqs) and all(not r.diverged for r in res)
    for r in reqs:
        alone = run_alone(r, SP_PRECISION_MIXED)
        w = _records_close(res[r.ident], alone, 1e-12)
        _log("flow %d %s: served vs alone %.2g, latency %.3f s" % (r.ident, r.params, w, res[r.ident].latency))
