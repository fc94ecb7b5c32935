"""The benchmark configurations at their stated sizes, through the C ABI,
against the FP64 CPU oracle.

* square lattice J = 1, range 4, 32 frequencies in [0.005, 50], cutoff 50 ->
  0.01 x 0.98 (the CPU-runnable configuration): computeStep at four cutoffs on
  states the GPU flow reached, and the whole FP64 correlation flow against the
  oracle's whole FP64 flow;
* mixed precision on simple cubic J1 = 1, J2 = 0.3, range 6, 64 frequencies in
  [0.002, 100]: sampled dGamma items at three cutoffs and every item at the
  middle one against the FP64 oracle (dSigma is FP64 in both forms), and the
  whole mixed correlation flow against the FP64 GPU flow;
* TRI kagome J = 1, D = 0.2 out of plane, range 5, 48 frequencies in
  [0.005, 50], on states the GPU flow reached: every dGamma item and dSigma at
  the first cutoff and a quarter into the schedule against the oracle's full
  evaluation; sampled dGamma items, dSigma and all ten correlations at three
  cutoffs;
* core options: SU(2) with spin S = 1 and with a normalization, TRI with a
  normalization.

Bars: FP64 1e-10 normwise relative error per output array, and elementwise
relative error with a floor of 1e-4 (derivatives) or 1e-6 (correlations) of
the array's largest entry (entries below the floor are held to the bar in
units of the floor); the mixed form
states its normwise and elementwise errors against the FP64 oracle and is held
to 1e-6 normwise.
"""
import sys

import numpy as np
import pytest

from helpers import cfg1_problem, cfg2_problem, cfg4_problem, elem_rel, exp_problem, max_rel, tri_problem_of

pytestmark = pytest.mark.gpu

TOL = 1e-10
TOL_MIXED = 1e-6
# elementwise floors (fractions of the array's largest entry): the derivative
# sums are reassociated on the GPU (lattice bubble as rank updates, reduction
# order), so an entry carries an absolute rounding error of a few 1e-16 of the
# array's largest entry; entries below 1e-4 of it are held to 1e-10 of the
# floor (1e-14 of the largest entry).  The correlations keep a floor of 1e-6.
FLOOR = 1e-4
FLOOR_CORR = 1e-6
# the whole-array checks meet the worst entries of millions: the terms of an
# entry can be far larger than the entry, so entries carry an absolute rounding
# error of up to a few 1e-14 of the array's largest entry; entries below 1e-3
# of it are held to 1e-13 of it
FLOOR_ALL = 1e-3


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(name, got, ref, tol, elementwise=True, floor=FLOOR):
    """Normwise bar, plus the elementwise bar for the FP64 form (the mixed
    form's FP32 rounding is relative to the array, so its small entries are
    reported, not held to the bar)."""
    e, el = max_rel(got, ref), elem_rel(got, ref, floor)
    _log("    %-16s normwise %.3g  elementwise %.3g" % (name, e, el))
    assert e < tol, name
    if elementwise:
        assert el < tol, name
    return e


def _states_along_flow(ctx_cls, prob, ks, precision=None):
    """The GPU flow's own states at the cutoff indices ks (FP64 unless given)."""
    from spinparser_amd.capi import SP_PRECISION_FP64
    ctx = ctx_cls(prob, SP_PRECISION_FP64 if precision is None else precision)
    ctx.set_state(float(prob.cutoffs[0]), *prob.initial_state(np.float64))
    out, pos = {}, 0
    for k in sorted(ks):
        if k > pos:
            _, div = ctx.run_flow(prob.cutoffs[pos:k + 1])
            assert not div
            pos = k
        out[k] = ctx.get_state()
    ctx.close()
    return out


def test_square_r4_nw32_steps():
    from oracle import Oracle
    from spinparser_amd.capi import FlowContext
    prob = cfg1_problem(nw=32)
    _log("square r4 nw32: L=%d nf=%d cutoffs=%d" % (prob.L, prob.n_freq, len(prob.cutoffs)))
    o = Oracle(prob)
    n = len(prob.cutoffs) - 1
    ks = [0, n // 3, 2 * n // 3, n - 1]
    states = _states_along_flow(FlowContext, prob, ks)
    ctx = FlowContext(prob)
    for k in ks:
        cut, v2, v4s, v4d = states[k]
        _log("  k=%d cutoff %.5g" % (k, cut))
        ref = o.compute_step(cut, v2, v4s, v4d)
        ctx.set_state(cut, v2, v4s, v4d)
        ctx.compute_step()
        got = ctx.get_flow()
        for name, g, r in zip(("dSigma", "dGamma spin", "dGamma density"), got, ref[:3]):
            _check(name, g, r, TOL)
    ctx.close()


@pytest.mark.timeout(600)
def test_square_r4_nw32_whole_flow():
    """The whole FP64 flow (421 evaluations) against the oracle's whole FP64 flow."""
    from oracle import Oracle
    from spinparser_amd.capi import FlowContext
    prob = cfg1_problem(nw=32)
    ctx = FlowContext(prob)
    ctx.set_state(float(prob.cutoffs[0]), *prob.initial_state(np.float64))
    recs, div = ctx.run_flow(prob.cutoffs)
    assert not div
    recs.append(ctx.measure_correlation())
    ctx.close()
    _log("square r4 nw32 whole flow: %d GPU records; oracle flow running" % len(recs))
    ozz, odd, on, _ = Oracle(prob).run_flow(np.float64)
    assert on == len(recs)
    worst = [0.0, 0.0]
    for k in range(on):
        for c, ref in ((0, ozz[k]), (1, odd[k])):
            if np.abs(ref).max() == 0.0:
                assert np.abs(recs[k][c]).max() == 0.0
                continue
            worst[0] = max(worst[0], max_rel(recs[k][c], ref))
            worst[1] = max(worst[1], elem_rel(recs[k][c], ref, FLOOR_CORR))
    _log("  correlations over %d records: normwise %.3g elementwise %.3g" % (on, worst[0], worst[1]))
    assert worst[0] < TOL and worst[1] < TOL


@pytest.mark.timeout(600)
def test_cubic_r6_nw64_mixed_vs_oracle():
    from oracle import Oracle
    from spinparser_amd.capi import SP_PRECISION_FP64, SP_PRECISION_MIXED, FlowContext
    prob = cfg2_problem()
    _log("cubic r6 nw64 mixed: L=%d nf=%d" % (prob.L, prob.n_freq))
    o = Oracle(prob)
    n = len(prob.cutoffs) - 1
    ks = [0, n // 2, n - 1]
    states = _states_along_flow(FlowContext, prob, ks)
    rng = np.random.default_rng(11)
    L = prob.L
    ctx = FlowContext(prob, SP_PRECISION_MIXED)
    worst = {"dSigma": 0.0, "dGamma spin": 0.0, "dGamma density": 0.0}
    for k in ks:
        cut, v2, v4s, v4d = states[k]
        _log("  k=%d cutoff %.5g" % (k, cut))
        ctx.set_state(cut, v2, v4s, v4d)
        ctx.compute_step()
        dv2, dss, ddd = ctx.get_flow()
        dv2_ref = o.dsigma(cut, v2, v4s, v4d)
        items = np.sort(rng.choice(prob.n_freq, 64, replace=False))
        gs, gd, rs, rd = [], [], [], []
        for it in items:
            a, b, _ = o.items(cut, v2, dv2_ref, v4s, v4d, int(it), int(it) + 1)
            sl = slice(it * L, (it + 1) * L)
            gs.append(dss[sl]); gd.append(ddd[sl]); rs.append(a[sl]); rd.append(b[sl])
        worst["dSigma"] = max(worst["dSigma"], _check("dSigma", dv2, dv2_ref, TOL))
        worst["dGamma spin"] = max(worst["dGamma spin"],
                                   _check("dGamma spin", np.concatenate(gs), np.concatenate(rs), TOL_MIXED,
                                          elementwise=False))
        worst["dGamma density"] = max(worst["dGamma density"],
                                      _check("dGamma density", np.concatenate(gd), np.concatenate(rd), TOL_MIXED,
                                             elementwise=False))
        if k == ks[1]:
            # every item (6.1 million values per array) against the oracle's full evaluation
            _, fs, fd, _ = o.compute_step(cut, v2, v4s, v4d)
            _check("dGamma spin (all)", dss, fs, TOL_MIXED, elementwise=False)
            _check("dGamma density (all)", ddd, fd, TOL_MIXED, elementwise=False)
    ctx.close()
    _log("  mixed vs FP64 oracle, max over 3 cutoffs x 64 items: %s" % worst)
    # the whole correlation flow, mixed against FP64 (655 records)
    flows = {}
    for prec in (SP_PRECISION_FP64, SP_PRECISION_MIXED):
        c = FlowContext(prob, prec)
        c.set_state(float(prob.cutoffs[0]), *prob.initial_state(np.float64))
        flows[prec], div = c.run_flow(prob.cutoffs)
        assert not div
        c.close()
    a, b = flows[SP_PRECISION_MIXED], flows[SP_PRECISION_FP64]
    assert len(a) == len(b) == n
    ezz = max(max_rel(x[0], y[0]) for x, y in zip(a, b) if np.abs(y[0]).max() > 0)
    edd = max(max_rel(x[1], y[1]) for x, y in zip(a, b) if np.abs(y[1]).max() > 0)
    _log("  correlations mixed vs FP64 over %d records: zz %.3g dd %.3g" % (n, ezz, edd))
    assert ezz < TOL_MIXED and edd < TOL_MIXED


@pytest.mark.timeout(600)
def test_kagome_dm_r5_nw48_sampled():
    from oracle import TRI_GROUPS, TriOracle
    from spinparser_amd.capi import TriFlowContext
    p = cfg4_problem()
    _log("TRI kagome-DM r5 nw48: L=%d O=%d nf=%d" % (p.L, p.overlap_total, p.n_freq))
    o = TriOracle(p)
    n = len(p.cutoffs) - 1
    ks = [0, n // 2, n - 1]
    full = {0, n // 4}  # every dGamma item (17.2 million values) against the oracle's full evaluation
    ctx = TriFlowContext(p)
    ctx.set_state(float(p.cutoffs[0]), *p.initial_state(np.float64))
    rng = np.random.default_rng(5)
    L16 = 16 * p.L
    for k in range(n):
        ctx.compute_step()
        if k in full:
            cut = float(p.cutoffs[k])
            _, v2, v4 = ctx.get_state()
            dv2, dv4 = ctx.get_flow()
            r2, r4, _ = o.compute_step(cut, v2, v4)
            _log("  k=%d cutoff %.5g, all items" % (k, cut))
            _check("dSigma", dv2, r2, TOL)
            _check("dGamma (all)", dv4, r4, TOL, floor=FLOOR_ALL)
        if k in ks:
            cut = float(p.cutoffs[k])
            _log("  k=%d cutoff %.5g" % (k, cut))
            _, v2, v4 = ctx.get_state()
            dv2, dv4 = ctx.get_flow()
            _check("dSigma", dv2, o.dsigma(cut, v2, v4), TOL)
            items = np.sort(rng.choice(p.n_freq, 32, replace=False))
            got, ref = [], []
            for it in items:
                r4, _ = o.items(cut, v2, dv2, v4, int(it), int(it) + 1, dtype=np.float64)
                sl = slice(it * L16, (it + 1) * L16)
                got.append(dv4[sl]); ref.append(r4[sl])
            _check("dGamma", np.concatenate(got), np.concatenate(ref), TOL)
            corr_g = ctx.measure_correlation()
            corr_o = o.correlation(cut, v2, v4, dtype=np.float64)
            for i, g in enumerate(TRI_GROUPS):
                if np.abs(corr_o[g]).max() > 0:
                    _check(g, corr_g[i], corr_o[g], TOL, floor=FLOOR_CORR)
        assert not ctx.is_diverged()
        ctx.finalize_step(float(p.cutoffs[k + 1]))
    ctx.close()


def test_core_options_spin_and_normalization():
    """spin = 1 and normalization for SU(2) (src/SU2/SU2FrgCore.cpp:19-29),
    normalization for TRI (src/TRI/TRIFrgCore.cpp:22-27): computeStep and the
    correlation against the oracle on the flow's own state."""
    from oracle import Oracle, TRI_GROUPS, TriOracle
    from spinparser_amd.capi import FlowContext, TriFlowContext
    cases = [("SU2 spin=1", exp_problem("square", 3, "square-heisenberg", {"j": 1.0}, 12, 0.005, 50.0, 50.0, 0.01,
                                         0.98, spin="1.0")),
             ("SU2 spin=1.5 normalization=0.5", exp_problem("square", 3, "square-heisenberg", {"j": 1.0}, 12, 0.005,
                                                             50.0, 50.0, 0.01, 0.98,
                                                             extra={"spin": "1.5", "normalization": "0.5"}))]
    for name, prob in cases:
        _log("%s: spin %.2f normalization %.2f" % (name, prob.spin, prob.normalization))
        assert prob.spin != 0.5
        o = Oracle(prob)
        for k in (0, 120):  # the spin-3/2, normalization-1/2 flow diverges after 161 records
            cut, v2, v4s, v4d = _states_along_flow(FlowContext, prob, [k])[k]
            ctx = FlowContext(prob)
            ctx.set_state(cut, v2, v4s, v4d)
            ctx.compute_step()
            got = ctx.get_flow()
            zz, dd = ctx.measure_correlation()
            ctx.close()
            ref = o.compute_step(cut, v2, v4s, v4d)
            for nm, g, r in zip(("dSigma", "dGamma spin", "dGamma density"), got, ref[:3]):
                if np.abs(r).max() > 0:
                    _check(nm, g, r, TOL)
            zr, dr = o.correlation(cut, v2, v4s, v4d)
            _check("corr zz", zz, zr, TOL, floor=FLOOR_CORR)
            _check("corr dd", dd, dr, TOL, floor=FLOOR_CORR)
    p = tri_problem_of("kagome", 2, "kagome-DM", {"j": 1.0, "d": 0.2}, 6, 0.05, 10.0, 10.0, 0.1, 0.9,
                       extra={"normalization": "2.0"})
    _log("TRI normalization %.2f" % p.normalization)
    assert p.normalization == 2.0
    o = TriOracle(p)
    cut = float(p.cutoffs[15])
    v2, v4 = p.initial_state(np.float64)
    ctx = TriFlowContext(p)
    ctx.set_state(cut, v2, v4)
    ctx.compute_step()
    dv2, dv4 = ctx.get_flow()
    corr = ctx.measure_correlation()
    ctx.close()
    r2, r4, _ = o.compute_step(cut, v2, v4, dtype=np.float64)
    _check("TRI dGamma", dv4, r4, TOL)
    co = o.correlation(cut, v2, v4, dtype=np.float64)
    for i, g in enumerate(TRI_GROUPS):
        if np.abs(co[g]).max() > 0:
            _check(g, corr[i], co[g], TOL, floor=FLOOR_CORR)
