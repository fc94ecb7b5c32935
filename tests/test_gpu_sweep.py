"""The served J2/J1 sweep at its stated sizes (simple cubic J1 = 1, J2 ~ U[0,1],
range 3..7, 32/48/64 frequencies in [0.002, 100], cutoff 100 -> 0.005 x 0.985,
PCG64 seed 2024): nine of its 64 flows, among them the ones whose pure
mixed-precision correlations strayed furthest from FP64 (flows 3, 60, 52, 45:
up to 6.6e-5), a range-7, 64-frequency flow (63) and a flow that diverges
(46, range 6: its FP64 flow stops after 420 records), served together in the
mixed form with the precision schedule (FP64 form at cutoffs up to three
times the coupling scale) under a memory budget that holds at most four.

* served = the same flow alone with the same schedule, to 1e-12 normwise on
  every record (continuous batching changes nothing);
* served vs the FP64 flow run alone: the same number of records and the same
  divergence verdict, every record within 1e-6 normwise (a diverging flow: up
  to ten records before its end);
* against the FP64 oracle: sampled dGamma items at one cutoff in the mixed
  part of the schedule (1e-6) and one in the FP64 part (1e-10), on states the
  scheduled flow reached.
"""
import sys

import numpy as np
import pytest

from helpers import max_rel

pytestmark = pytest.mark.gpu

FLOWS = (3, 60, 52, 45, 63, 0, 5, 22, 46)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def requests():
    import dataclasses
    from spinparser_amd.serve import sweep_requests
    reqs = sweep_requests(64, seed=2024, rate_per_s=1.0)  # the draws of the benchmark sweep
    return [dataclasses.replace(reqs[i], arrival=0.0) for i in FLOWS]


@pytest.mark.timeout(900)
def test_served_sweep_schedule_vs_fp64(requests):
    from spinparser_amd.capi import SP_PRECISION_FP64, SP_PRECISION_MIXED, footprint_bytes
    from spinparser_amd.serve import PrecisionSchedule, SweepServer, run_alone
    sched = PrecisionSchedule()
    assert any(r.params["range"] == 7 and r.params["nw"] == 64 for r in requests)
    fps = sorted(footprint_bytes(r.problem, SP_PRECISION_MIXED) for r in requests)
    budget = int(sum(fps[-4:]))  # at most four flows resident
    srv = SweepServer(budget_bytes=budget, max_flows=8, precision=SP_PRECISION_MIXED, schedule=sched)
    served = {r.ident: r for r in srv.run(requests, log=_log)}
    srv.close()
    assert srv.peak_bytes <= budget
    _log("served %d flows, peak %d resident, %d mixed / %d FP64 evaluations"
         % (len(served), srv.peak_flows, srv.mixed_steps, srv.fp64_steps))
    assert srv.mixed_steps > 0 and srv.fp64_steps > 0
    for r in requests:
        s = served[r.ident]
        alone = run_alone(r, SP_PRECISION_MIXED, schedule=sched)
        assert len(alone.zz) == len(s.zz) and alone.diverged == s.diverged
        w_alone = max(max(max_rel(a, b) for a, b in zip(s.zz, alone.zz)),
                      max(max_rel(a, b) for a, b in zip(s.dd, alone.dd)))
        f64 = run_alone(r, SP_PRECISION_FP64, queued=True)
        assert len(f64.zz) == len(s.zz), (r.ident, len(f64.zz), len(s.zz))
        assert f64.diverged == s.diverged
        # a diverging flow's last records are dominated by the blow-up; its
        # records up to ten before the divergence are compared
        n = len(s.zz) - 10 if s.diverged else len(s.zz)
        w64 = max(max(max_rel(a, b) for a, b in zip(s.zz[:n], f64.zz[:n])),
                  max(max_rel(a, b) for a, b in zip(s.dd[:n], f64.dd[:n])))
        _log("flow %d %s: %d records diverged %s; served vs alone %.2g, vs FP64 %.2g"
             % (r.ident, r.params, n, s.diverged, w_alone, w64))
        assert w_alone <= 1e-12
        assert w64 < 1e-6


def test_scheduled_flow_vs_oracle(requests):
    from oracle import Oracle
    from spinparser_amd.capi import SP_PRECISION_FP64, SP_PRECISION_MIXED, FlowContext
    from spinparser_amd.serve import PrecisionSchedule, coupling_scale
    sched = PrecisionSchedule()
    req = requests[0]  # flow 3: range 5, 48 frequencies
    prob = req.problem
    o = Oracle(prob)
    J = coupling_scale(prob)
    cut = np.asarray(prob.cutoffs, dtype=np.float64)
    k_mixed = int(np.nonzero(cut > 10.0 * J)[0][-1])
    k_fp64 = int(np.nonzero(cut <= 0.5 * J)[0][0])
    ctx = FlowContext(prob, SP_PRECISION_MIXED)
    ctx.set_state(float(cut[0]), *prob.initial_state(np.float64))
    rng = np.random.default_rng(3)
    L = prob.L
    for k in range(k_fp64 + 1):
        want = sched.precision(prob, float(cut[k]))
        if want != ctx.precision:
            ctx.set_precision(want)
        ctx.compute_step()
        if k in (k_mixed, k_fp64):
            _, v2, v4s, v4d = ctx.get_state()
            dv2, dss, ddd = ctx.get_flow()
            dv2_ref = o.dsigma(float(cut[k]), v2, v4s, v4d)
            got, ref = [], []
            for it in np.sort(rng.choice(prob.n_freq, 64, replace=False)):
                a, b, _ = o.items(float(cut[k]), v2, dv2_ref, v4s, v4d, int(it), int(it) + 1)
                sl = slice(it * L, (it + 1) * L)
                got += [dss[sl], ddd[sl]]
                ref += [a[sl], b[sl]]
            e_sig = max_rel(dv2, dv2_ref)
            e = max_rel(np.concatenate(got), np.concatenate(ref))
            form = "mixed" if want == SP_PRECISION_MIXED else "FP64"
            _log("flow %d at cutoff %.4g (%s form): dSigma %.3g, 64 dGamma items %.3g" % (req.ident, cut[k], form,
                                                                                         e_sig, e))
            assert e_sig < 1e-10
            assert e < (1e-6 if want == SP_PRECISION_MIXED else 1e-10)
        ctx.finalize_step(float(cut[k + 1]))
    ctx.close()
