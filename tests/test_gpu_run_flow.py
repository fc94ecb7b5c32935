"""sp_run_flow, the queued cutoff loop, against the step-by-step ABI calls.

The same kernels run in the same order on the same data, so every record, the
final state and the final flow must agree bit for bit: on a flow that reaches
its last cutoff and on one that diverges mid-way (the loop must stop at the
same step and leave the context as the step-by-step loop does).
"""
import sys

import numpy as np
import pytest

from helpers import cfg1_problem, exp_problem

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _step_by_step(ctx, cutoffs):
    """SpinParser::runCore with one ABI call per stage (src/SpinParser.cpp:141-172)."""
    recs = []
    for k in range(len(cutoffs) - 1):
        ctx.compute_step()
        recs.append(ctx.measure_correlation())
        if ctx.is_diverged():
            return recs, True
        ctx.finalize_step(float(cutoffs[k + 1]))
    return recs, False


def _compare(prob, precision, expect_diverged):
    from spinparser_amd.capi import FlowContext
    init = prob.initial_state(np.float64)
    a = FlowContext(prob, precision)
    b = FlowContext(prob, precision)
    for c in (a, b):
        c.set_state(float(prob.cutoffs[0]), *init)
    ra, da = _step_by_step(a, prob.cutoffs)
    rb, db = b.run_flow(prob.cutoffs)
    _log("L=%d nw=%d: %d records step by step, %d queued, diverged %s / %s" % (prob.L, prob.nw, len(ra), len(rb),
                                                                              da, db))
    assert da == db == expect_diverged
    assert len(ra) == len(rb)
    for (za, xa), (zb, xb) in zip(ra, rb):
        assert np.array_equal(za, zb) and np.array_equal(xa, xb)
    sa, sb = a.get_state(), b.get_state()
    assert sa[0] == sb[0]  # cutoff
    for x, y in zip(sa[1:], sb[1:]):
        assert np.array_equal(x, y, equal_nan=True)
    if expect_diverged:
        assert b.is_diverged()
        for x, y in zip(a.get_flow(), b.get_flow()):
            assert np.array_equal(x, y, equal_nan=True)
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_flow_equals_steps(precision):
    from spinparser_amd.capi import SP_PRECISION_FP64, SP_PRECISION_MIXED
    prec = SP_PRECISION_FP64 if precision == "fp64" else SP_PRECISION_MIXED
    _compare(cfg1_problem(nw=8), prec, False)


def test_run_flow_diverging():
    """Simple cubic J1-J2 at J2 = 0.0385 (range 6, N_w = 32): the flow breaks
    down before the last cutoff (a NaN in dGamma)."""
    from spinparser_amd.capi import SP_PRECISION_FP64
    prob = exp_problem("cubic", 6, "cubic-j1j2heisenberg", {"j1": 1.0, "j2": 0.03845590068527294}, 32, 0.002, 100.0,
                       100.0, 0.005, 0.985)
    _compare(prob, SP_PRECISION_FP64, True)
