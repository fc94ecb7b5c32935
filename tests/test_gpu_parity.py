"""GPU FP64 flow vs the FP64 CPU oracle, through the C ABI.

Tolerance: normwise max-relative error 1e-10 per output array (FP64 form);
frequency arguments, window predicates and mesh brackets are evaluated with the
same IEEE operations on both sides, only summation order differs.
"""
import sys

import numpy as np
import pytest

from helpers import cfg1_problem, cfg2_problem, elem_rel, max_rel, ref1_problem, python_obs_problem
from oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
FLOOR_ALL = 1e-3  # elementwise floor of the whole-array checks (fraction of the largest entry)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx(prob):
    from spinparser_amd.capi import FlowContext
    return FlowContext(prob)


def _advance_oracle(o, prob, k):
    """State after k Euler steps of the FP64 oracle."""
    v2, v4s, v4d = prob.initial_state(np.float64)
    for i in range(k):
        dv2, dss, ddd, _ = o.compute_step(prob.cutoffs[i], v2, v4s, v4d)
        step = float(prob.cutoffs[i + 1]) - float(prob.cutoffs[i])
        v2 = v2 + step * dv2
        v4s = v4s + step * dss
        v4d = v4d + step * ddd
    return v2, v4s, v4d


@pytest.mark.parametrize("case", ["ref1", "square_nw8", "square_nw16", "honeycomb"])
def test_compute_step_matches_oracle(case):
    prob = {"ref1": ref1_problem, "square_nw8": lambda: cfg1_problem(nw=8),
            "square_nw16": lambda: cfg1_problem(nw=16), "honeycomb": python_obs_problem}[case]()
    _log("case %s L=%d nw=%d" % (case, prob.L, prob.nw))
    o = Oracle(prob)
    ctx = _ctx(prob)
    ks = sorted({0, len(prob.cutoffs) // 3, 2 * len(prob.cutoffs) // 3, len(prob.cutoffs) - 2})
    state = prob.initial_state(np.float64)
    done = 0
    for k in ks:
        # advance the oracle state to cutoff index k
        while done < k:
            dv2, dss, ddd, _ = o.compute_step(prob.cutoffs[done], *state)
            step = float(prob.cutoffs[done + 1]) - float(prob.cutoffs[done])
            state = (state[0] + step * dv2, state[1] + step * dss, state[2] + step * ddd)
            done += 1
        cut = float(prob.cutoffs[k])
        ref = o.compute_step(cut, *state)
        ctx.set_state(cut, *state)
        ctx.compute_step()
        dv2, dss, ddd = ctx.get_flow()
        e = (max_rel(dv2, ref[0]), max_rel(dss, ref[1]), max_rel(ddd, ref[2]))
        _log("  k=%d cutoff=%.6g err dSigma %.3g dGs %.3g dGd %.3g" % (k, cut, *e))
        assert e[0] < TOL and e[1] < TOL and e[2] < TOL
        assert not ctx.is_diverged()
    ctx.close()


def test_correlation_matches_oracle():
    prob = ref1_problem()
    o = Oracle(prob)
    state = _advance_oracle(o, prob, 20)
    cut = float(prob.cutoffs[20])
    zz_ref, dd_ref = o.correlation(cut, *state)
    ctx = _ctx(prob)
    ctx.set_state(cut, *state)
    zz, dd = ctx.measure_correlation()
    _log("correlation err zz %.3g dd %.3g" % (max_rel(zz, zz_ref), max_rel(dd, dd_ref)))
    assert max_rel(zz, zz_ref) < TOL and max_rel(dd, dd_ref) < TOL
    ctx.close()


def test_euler_update_matches():
    prob = cfg1_problem(nw=8)
    o = Oracle(prob)
    state = _advance_oracle(o, prob, 30)
    cut, new = float(prob.cutoffs[30]), float(prob.cutoffs[31])
    ctx = _ctx(prob)
    ctx.set_state(cut, *state)
    ctx.compute_step()
    dv2, dss, ddd = ctx.get_flow()
    ctx.finalize_step(new)
    c, v2, v4s, v4d = ctx.get_state()
    assert c == new
    step = new - cut
    assert max_rel(v4s, state[1] + step * dss) < 1e-14
    assert max_rel(v4d, state[2] + step * ddd) < 1e-14
    assert max_rel(v2, state[0] + step * dv2) < 1e-14
    ctx.close()


def test_full_flow_matches_golden_fp32():
    """The FP64 GPU flow reproduces the reference's float golden file to 1e-5."""
    import os
    from helpers import GOLDEN
    from spinparser_amd.frgcore import FrgCoreFactory, run_core
    from spinparser_amd.task import MeasurementSpec
    golden = np.load(os.path.join(GOLDEN, "reference1_su2.npz"))
    prob = ref1_problem()
    core = FrgCoreFactory.newFrgCore("SU2", prob, [MeasurementSpec("correlation")], {}, backend="hip")
    assert run_core(core, prob.cutoffs)
    recs = core.measurements()[0].records
    assert len(recs) == len(prob.cutoffs)
    worst = 0.0
    for k, r in enumerate(recs):
        worst = max(worst, np.abs(r.zz - golden["SU2CorZZ/data/measurement_%d/data" % k]).max(),
                    np.abs(r.dd - golden["SU2CorDD/data/measurement_%d/data" % k]).max())
    _log("full flow vs golden worst abs %.3g" % worst)
    assert worst < 1e-5
    core.close()


def test_divergence_flag_from_host_nan():
    """A NaN planted in the state by the host is reported by isDiverged()."""
    prob = cfg1_problem(nw=8)
    ctx = _ctx(prob)
    v2, v4s, v4d = prob.initial_state(np.float64)
    ctx.set_state(float(prob.cutoffs[40]), v2, v4s, v4d)
    ctx.compute_step()
    assert not ctx.is_diverged()
    v4s = v4s.copy()
    v4s[:] = np.nan
    ctx.set_state(float(prob.cutoffs[40]), v2, v4s, v4d)
    ctx.compute_step()
    assert ctx.is_diverged()
    ctx.close()


def test_cfg2_sampled_items():
    """Simple cubic J1-J2 range 6, N_w = 64 (the benchmark flow) on states the
    GPU flow itself reached: dSigma and a sample of 96 dGamma items at the
    first cutoff, and dSigma and every one of the 133,120 items (6.1 million
    values per array) at cutoffs 300 and 560 against the oracle's full
    evaluation, normwise and elementwise with a floor of 1e-3 of the array's
    largest entry.  The GPU reassociates the sums (lattice bubble as rank
    updates, reduction order), and the terms of an entry can be far larger than
    the entry, so every entry carries an absolute rounding error of up to a few
    1e-14 of the array's largest entry (4.2e-14 at cutoff 560); entries below
    the floor are held to 1e-13 of the largest entry."""
    prob = cfg2_problem()
    _log("cfg2 L=%d nw=%d nf=%d" % (prob.L, prob.nw, prob.n_freq))
    o = Oracle(prob)
    ctx = _ctx(prob)
    v2, v4s, v4d = prob.initial_state(np.float64)
    ctx.set_state(float(prob.cutoffs[0]), v2, v4s, v4d)
    rng = np.random.default_rng(7)
    sampled, full = {0}, {300, 560}
    L = prob.L
    for k in range(561):
        ctx.compute_step()
        if k in sampled | full:
            cut = float(prob.cutoffs[k])
            _, s2, s4s, s4d = ctx.get_state()
            dv2, dss, ddd = ctx.get_flow()
            if k in full:
                dv2_ref, rs, rd, _ = o.compute_step(cut, s2, s4s, s4d)
                e_sig = max_rel(dv2, dv2_ref)
                e_s, e_d = max_rel(dss, rs), max_rel(ddd, rd)
                el_s, el_d = elem_rel(dss, rs, FLOOR_ALL), elem_rel(ddd, rd, FLOOR_ALL)
                _log("  k=%d cutoff=%.5g dSigma err %.3g, all items: spin %.3g (elementwise %.3g), "
                     "density %.3g (elementwise %.3g)" % (k, cut, e_sig, e_s, el_s, e_d, el_d))
                assert max(e_sig, e_s, e_d, el_s, el_d) < TOL
            else:
                dv2_ref = o.dsigma(cut, s2, s4s, s4d)
                items = np.sort(rng.choice(prob.n_freq, 96, replace=False))
                got, ref = [], []
                for it in items:
                    rs, rd, _ = o.items(cut, s2, dv2_ref, s4s, s4d, int(it), int(it) + 1)
                    sl = slice(it * L, (it + 1) * L)
                    got += [dss[sl], ddd[sl]]
                    ref += [rs[sl], rd[sl]]
                e_sig = max_rel(dv2, dv2_ref)
                e_items = max_rel(np.concatenate(got), np.concatenate(ref))
                _log("  k=%d cutoff=%.5g dSigma err %.3g sampled dGamma err %.3g" % (k, cut, e_sig, e_items))
                assert e_sig < TOL and e_items < TOL
        ctx.finalize_step(float(prob.cutoffs[k + 1]))
    ctx.close()
