"""GPU parity of the XYZ core (HIP path through the C ABI) against the FP64 CPU oracle.

Bar: 1e-10 normwise relative error (max|gpu - oracle| / max|oracle|) on dSigma,
dGamma and the four correlation outputs; the full golden flow additionally within
the reference's own 1e-5 absolute tolerance of its float golden file
(test/scripted/assets/test_eval.py:14,28-29).

Cases: the reference's XYZ golden problem (hyperhoneycomb Heisenberg-Kitaev,
range 3, 14 frequencies; x <-> y spin permutations on inverted sites and
overlaps) at oracle-advanced states, and the honeycomb Kitaev-Gamma task run
with the XYZ core (its off-diagonal Gamma couplings are ignored by the XYZ
initial condition, as in the reference; all six spin permutations occur among
its overlap transformations) at frequency-dependent states.  Both dGamma
kernels are covered: the default one (lattice bubble as rank updates of
per-pair matrices in LDS) and the direct one used when those matrices exceed
the LDS of a workgroup (forced here with SPFRG_XYZ_BUBBLE=direct).
"""
import sys

import numpy as np
import pytest

from helpers import GOLDEN, GOLDEN_RES, manual_task_xml, max_rel, xyz_ref2_problem

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ref2():
    from oracle import XyzOracle
    p = xyz_ref2_problem()
    return p, XyzOracle(p)


def _advanced_state(o, p, k):
    """FP64 oracle state after k Euler steps from the initial condition."""
    if k == 0:
        v2, v4 = p.initial_state(np.float64)
        return float(p.cutoffs[0]), v2, v4
    _, n, (v2, v4) = o.run_flow(dtype=np.float64, nthreads=8, ncut=k + 1)
    assert n == k + 1
    return float(p.cutoffs[k]), v2, v4


def _compare(p, o, cut, v2, v4, label):
    from spinparser_amd.capi import XyzFlowContext
    from oracle import XYZ_GROUPS
    _log("%s at cutoff %.4g" % (label, cut))
    dv2_o, dv4_o, _ = o.compute_step(cut, v2, v4, dtype=np.float64, nthreads=8)
    ctx = XyzFlowContext(p)
    ctx.set_state(cut, v2, v4)
    ctx.compute_step()
    dv2_g, dv4_g = ctx.get_flow()
    assert not ctx.is_diverged()
    corr_g = ctx.measure_correlation()
    ctx.close()
    corr_o = o.correlation(cut, v2, v4, dtype=np.float64)
    e2, e4 = max_rel(dv2_g, dv2_o), max_rel(dv4_g, dv4_o)
    ec = max(max_rel(corr_g[i], corr_o[g]) for i, g in enumerate(XYZ_GROUPS))
    _log("  dSigma %.3g  dGamma %.3g  correlations %.3g" % (e2, e4, ec))
    assert np.abs(dv4_o).max() > 0.0
    assert e2 < TOL and e4 < TOL and ec < TOL


@pytest.mark.parametrize("k", [0, 12, 25, 33])
def test_xyz_step_parity_golden_problem(ref2, k):
    p, o = ref2
    cut, v2, v4 = _advanced_state(o, p, k)
    _compare(p, o, cut, v2, v4, "XYZ step parity: hyperhoneycomb-kitaev r3 nw=%d L=%d index %d" % (p.nw, p.L, k))


def test_xyz_full_flow_golden(ref2):
    """Whole flow on the GPU: every record within 1e-10 of the FP64 oracle flow
    and within the reference's 1e-5 of its float golden file."""
    from spinparser_amd.capi import XyzFlowContext
    from oracle import XYZ_GROUPS
    p, o = ref2
    rec_o, n_o, _ = o.run_flow(dtype=np.float64, nthreads=8)
    ctx = XyzFlowContext(p)
    v2, v4 = p.initial_state(np.float64)
    ctx.set_state(float(p.cutoffs[0]), v2, v4)
    recs = []
    for k in range(len(p.cutoffs) - 1):
        ctx.compute_step()
        recs.append(ctx.measure_correlation())
        assert not ctx.is_diverged()
        ctx.finalize_step(float(p.cutoffs[k + 1]))
    recs.append(ctx.measure_correlation())
    ctx.close()
    assert len(recs) == n_o
    g = np.load(GOLDEN + "/reference2_xyz.npz")
    worst_o = worst_g = 0.0
    for k in range(n_o):
        worst_o = max(worst_o, max_rel(recs[k], rec_o[k]))
        for i, grp in enumerate(XYZ_GROUPS):
            worst_g = max(worst_g, float(np.abs(recs[k][i] - g["%s/data/measurement_%d/data" % (grp, k)]).max()))
    _log("XYZ full flow: %d records, vs FP64 oracle %.3g, vs float golden %.3g (abs)" % (n_o, worst_o, worst_g))
    assert worst_o < TOL
    assert worst_g < 1e-5


@pytest.mark.parametrize("bubble", ["rank-update", "direct"])
def test_xyz_permuted_honeycomb_parity(bubble, monkeypatch):
    """Honeycomb Kitaev-Gamma range 3 with the XYZ core: every spin permutation
    among the overlap transformations; frequency-dependent states."""
    if bubble == "direct":
        monkeypatch.setenv("SPFRG_XYZ_BUBBLE", "direct")
    from spinparser_amd.problem import xyz_problem
    from spinparser_amd.task import parse_task
    from oracle import XyzOracle
    xml = manual_task_xml("honeycomb", 3, "honeycomb-kitaev-gamma", {"j": "0.2", "k": "1.0", "g": "-0.1"},
                          symmetry="XYZ")
    p = xyz_problem(parse_task(xml, True), GOLDEN_RES)
    assert len(np.unique(p.ovl_perm1, axis=0)) == 6
    o = XyzOracle(p)
    for k in (0, 10, 30):
        cut, v2, v4 = _advanced_state(o, p, min(k, 20))
        cut = float(p.cutoffs[k])
        _compare(p, o, cut, v2, v4, "XYZ honeycomb-kitaev-gamma r3 nw=%d L=%d O=%d (%s)"
                 % (p.nw, p.L, p.overlap_total, bubble))


@pytest.mark.timeout(600)
def test_xyz_honeycomb_kitaev_r8_nw64_along_flow():
    """The XYZ core at its benchmark size (honeycomb Heisenberg-Kitaev J = 0.1,
    K = -1, range 8: L = 23; 64 frequencies in [0.005, 50], cutoff 50 -> 0.01
    x 0.98) on states its own GPU flow reached: every dGamma item, dSigma and
    the four correlations at three cutoffs against the oracle's full
    evaluation, normwise, and dGamma elementwise with a floor of 1e-3 of the
    largest entry (whole arrays of 12 million entries: the reassociated sums
    leave the worst entries an absolute error of a few 1e-14 of it)."""
    import bench
    from helpers import elem_rel
    from oracle import XYZ_GROUPS, XyzOracle
    from spinparser_amd.capi import XyzFlowContext
    p = bench.build_core_problem("xyz")
    _log("XYZ honeycomb-kitaev r8 nw64: L=%d O=%d nf=%d" % (p.L, p.overlap_total, p.n_freq))
    o = XyzOracle(p)
    n = len(p.cutoffs) - 1
    ks = {0, n // 4, n // 2}
    ctx = XyzFlowContext(p)
    ctx.set_state(float(p.cutoffs[0]), *p.initial_state(np.float64))
    for k in range(n // 2 + 1):
        ctx.compute_step()
        if k in ks:
            cut = float(p.cutoffs[k])
            _, v2, v4 = ctx.get_state()
            dv2_g, dv4_g = ctx.get_flow()
            dv2_o, dv4_o, _ = o.compute_step(cut, v2, v4, dtype=np.float64)
            corr_g = ctx.measure_correlation()
            corr_o = o.correlation(cut, v2, v4, dtype=np.float64)
            e2, e4, el4 = max_rel(dv2_g, dv2_o), max_rel(dv4_g, dv4_o), elem_rel(dv4_g, dv4_o, 1e-3)
            ec = max(max_rel(corr_g[i], corr_o[g]) for i, g in enumerate(XYZ_GROUPS) if np.abs(corr_o[g]).max() > 0)
            _log("  k=%d cutoff %.5g: dSigma %.3g  dGamma %.3g (elementwise %.3g)  correlations %.3g"
                 % (k, cut, e2, e4, el4, ec))
            assert e2 < TOL and e4 < TOL and el4 < TOL and ec < TOL
        assert not ctx.is_diverged()
        ctx.finalize_step(float(p.cutoffs[k + 1]))
    ctx.close()
