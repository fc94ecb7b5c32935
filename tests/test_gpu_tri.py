"""GPU parity of the TRI core (HIP path through the C ABI) against the FP64 CPU oracle.

Bar: 1e-10 normwise relative error (max|gpu - oracle| / max|oracle|) on dSigma,
dGamma and all 10 correlation outputs; the full golden flow additionally within
the reference's own 1e-5 absolute tolerance of its float golden file
(test/scripted/assets/test_eval.py:14,28-29).

Cases: the reference's TRI golden problem (honeycomb Heisenberg-Kitaev-Gamma,
range 3, 14 frequencies) at oracle-advanced states, and the kagome lattice with
out-of-plane Dzyaloshinskii-Moriya coupling D = 0.2 at range 5 (L = 19, 771
overlap terms, spin permutations on half of the site transformations) on a
reduced frequency mesh.
"""
import sys

import numpy as np
import pytest

from helpers import GOLDEN, cfg4_problem, max_rel, tri_ref3_problem

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ref3():
    from oracle import TriOracle
    p = tri_ref3_problem()
    return p, TriOracle(p)


def _advanced_state(o, p, k):
    """FP64 oracle state after k Euler steps from the initial condition."""
    if k == 0:
        v2, v4 = p.initial_state(np.float64)
        return float(p.cutoffs[0]), v2, v4
    _, n, (v2, v4) = o.run_flow(dtype=np.float64, nthreads=8, ncut=k + 1)
    assert n == k + 1
    return float(p.cutoffs[k]), v2, v4


@pytest.mark.parametrize("k", [0, 12, 25, 33])
def test_tri_step_parity_golden_problem(ref3, k):
    from spinparser_amd.capi import TriFlowContext
    p, o = ref3
    cut, v2, v4 = _advanced_state(o, p, k)
    _log("TRI step parity: honeycomb-kitaev-gamma r3 nw=%d L=%d at cutoff index %d (%.4g)" % (p.nw, p.L, k, cut))
    dv2_o, dv4_o, _ = o.compute_step(cut, v2, v4, dtype=np.float64, nthreads=8)
    ctx = TriFlowContext(p)
    ctx.set_state(cut, v2, v4)
    ctx.compute_step()
    dv2_g, dv4_g = ctx.get_flow()
    assert not ctx.is_diverged()
    e2, e4 = max_rel(dv2_g, dv2_o), max_rel(dv4_g, dv4_o)
    _log("  dSigma %.3g  dGamma %.3g" % (e2, e4))
    assert e2 < TOL and e4 < TOL
    corr_g = ctx.measure_correlation()
    corr_o = o.correlation(cut, v2, v4, dtype=np.float64)
    from oracle import TRI_GROUPS
    ec = max(max_rel(corr_g[i], corr_o[g]) for i, g in enumerate(TRI_GROUPS))
    _log("  correlations %.3g" % ec)
    assert ec < TOL
    ctx.close()


def test_tri_full_flow_golden(ref3):
    """Whole flow on the GPU: every record within 1e-10 of the FP64 oracle flow
    and within the reference's 1e-5 of its float golden file."""
    from spinparser_amd.capi import TriFlowContext
    from oracle import TRI_GROUPS
    p, o = ref3
    rec_o, n_o, _ = o.run_flow(dtype=np.float64, nthreads=8)
    ctx = TriFlowContext(p)
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
    g = np.load(GOLDEN + "/reference3_tri.npz")
    worst_o = worst_g = 0.0
    for k in range(n_o):
        worst_o = max(worst_o, max_rel(recs[k], rec_o[k]))
        for i, grp in enumerate(TRI_GROUPS):
            worst_g = max(worst_g, float(np.abs(recs[k][i] - g["%s/data/measurement_%d/data" % (grp, k)]).max()))
    _log("TRI full flow: %d records, vs FP64 oracle %.3g, vs float golden %.3g (abs)" % (n_o, worst_o, worst_g))
    assert worst_o < TOL
    assert worst_g < 1e-5


def test_tri_kagome_dm_parity():
    """Kagome DM range 5 (spin-permuted site transformations) on an 8-point mesh."""
    from spinparser_amd.capi import TriFlowContext
    from oracle import TRI_GROUPS, TriOracle
    p = cfg4_problem(nw=8, lattice_range=5)
    o = TriOracle(p)
    for k in (0, 150, 300):
        cut, v2, v4 = _advanced_state(o, p, 0)
        cut = float(p.cutoffs[k])
        if k > 0:
            # a frequency-dependent state: one oracle Euler step of the initial condition
            dv2, dv4, _ = o.compute_step(cut, v2, v4, dtype=np.float64, nthreads=8)
            v2, v4 = v2 - 0.05 * cut * dv2, v4 - 0.05 * cut * dv4
        _log("TRI kagome-DM r5 nw=%d L=%d O=%d at cutoff %.4g" % (p.nw, p.L, p.overlap_total, cut))
        dv2_o, dv4_o, _ = o.compute_step(cut, v2, v4, dtype=np.float64, nthreads=8)
        ctx = TriFlowContext(p)
        ctx.set_state(cut, v2, v4)
        ctx.compute_step()
        dv2_g, dv4_g = ctx.get_flow()
        e2, e4 = max_rel(dv2_g, dv2_o), max_rel(dv4_g, dv4_o)
        corr_g = ctx.measure_correlation()
        corr_o = o.correlation(cut, v2, v4, dtype=np.float64)
        ec = max(max_rel(corr_g[i], corr_o[g]) for i, g in enumerate(TRI_GROUPS))
        _log("  dSigma %.3g  dGamma %.3g  correlations %.3g" % (e2, e4, ec))
        ctx.close()
        assert e2 < TOL and e4 < TOL and ec < TOL


def _gpu_flow(p):
    from spinparser_amd.capi import TriFlowContext
    ctx = TriFlowContext(p)
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
    return np.array(recs)


def test_tri_dm_flow_spin_symmetries():
    """Kagome J = 1 with out-of-plane DM D = +-0.3, range 3, 8 frequencies, whole flows on the GPU.

    The kernel's own records must respect the spin symmetries of the model
    (rotations about z: XX = YY, XY = -YX, XZ = YZ = ZX = ZY = 0; the pi rotation
    about x maps D onto -D: XY, YX, XZ, ZX odd in D, the rest even), and agree with
    the FP64 oracle's flow within the 1e-10 bar."""
    from helpers import tri_problem_of
    from oracle import TRI_GROUPS, TriOracle
    flows = {}
    for d in (0.3, -0.3):
        p = tri_problem_of("kagome", 3, "kagome-DM", {"j": 1.0, "d": d}, 8, 0.05, 20.0, 10.0, 0.8, 0.9)
        _log("TRI kagome-DM r3 D=%g L=%d O=%d, %d cutoffs" % (d, p.L, p.overlap_total, len(p.cutoffs)))
        rec = _gpu_flow(p)
        rec_o, n_o, _ = TriOracle(p).run_flow(dtype=np.float64, nthreads=8)
        assert rec.shape[0] == n_o
        worst = max(max_rel(rec[k], rec_o[k]) for k in range(n_o))
        _log("  vs FP64 oracle %.3g" % worst)
        assert worst < TOL
        flows[d] = rec
    g = {name: i for i, name in enumerate(TRI_GROUPS)}
    a, b = flows[0.3], flows[-0.3]
    tol = 1e-12 * np.abs(a).max()
    assert np.abs(a[:, g["TRICorXY"]]).max() > 1e-3 * np.abs(a).max()
    assert np.abs(a[:, g["TRICorXX"]] - a[:, g["TRICorYY"]]).max() < tol
    assert np.abs(a[:, g["TRICorXY"]] + a[:, g["TRICorYX"]]).max() < tol
    for name in ("TRICorXZ", "TRICorYZ", "TRICorZX", "TRICorZY"):
        assert np.abs(a[:, g[name]]).max() < tol, name
    parity = {"TRICorDD": 1, "TRICorXX": 1, "TRICorXY": -1, "TRICorXZ": -1, "TRICorYX": -1,
              "TRICorYY": 1, "TRICorYZ": 1, "TRICorZX": -1, "TRICorZY": 1, "TRICorZZ": 1}
    for name, sgn in parity.items():
        assert np.abs(a[:, g[name]] - sgn * b[:, g[name]]).max() < tol, name


def test_tri_large_lattice_workgroup_shape():
    """Kagome-DM range 7 (L = 34): above 32 sites the dGamma kernel runs in
    256-thread workgroups (six points per chunk at most) instead of 128; every
    dGamma item, dSigma and the ten correlations on the GPU flow's own states at
    three cutoffs against the oracle's full evaluation."""
    from helpers import elem_rel, tri_problem_of
    from oracle import TRI_GROUPS, TriOracle
    from spinparser_amd.capi import TriFlowContext
    p = tri_problem_of("kagome", 7, "kagome-DM", {"j": 1.0, "d": 0.2}, 8, 0.005, 50.0, 50.0, 0.01, 0.98)
    assert p.L > 32
    _log("TRI kagome-DM r7 nw8: L=%d O=%d" % (p.L, p.overlap_total))
    o = TriOracle(p)
    n = len(p.cutoffs) - 1
    ks = {0, n // 2, n - 1}
    ctx = TriFlowContext(p)
    ctx.set_state(float(p.cutoffs[0]), *p.initial_state(np.float64))
    for k in range(n):
        ctx.compute_step()
        if k in ks:
            cut = float(p.cutoffs[k])
            _, v2, v4 = ctx.get_state()
            dv2, dv4 = ctx.get_flow()
            r2, r4, _ = o.compute_step(cut, v2, v4)
            corr_g = ctx.measure_correlation()
            corr_o = o.correlation(cut, v2, v4, dtype=np.float64)
            e2, e4, el4 = max_rel(dv2, r2), max_rel(dv4, r4), elem_rel(dv4, r4, 1e-3)
            ec = max(max_rel(corr_g[i], corr_o[g]) for i, g in enumerate(TRI_GROUPS) if np.abs(corr_o[g]).max() > 0)
            _log("  k=%d cutoff %.4g: dSigma %.3g  dGamma %.3g (elementwise %.3g)  correlations %.3g"
                 % (k, cut, e2, e4, el4, ec))
            assert e2 < TOL and e4 < TOL and el4 < TOL and ec < TOL
        assert not ctx.is_diverged()
        ctx.finalize_step(float(p.cutoffs[k + 1]))
    ctx.close()


# one lattice per workgroup shape of the 128-thread dGamma kernel: compiled
# (site stride, points per chunk) = (L rounded up to 4, 64 // L) for L = 13..32
# (L = 21 excepted), and the run-time form below 13 sites
SHAPE_CASES = [
    ("kagome", 3, "kagome-DM", {"j": 1.0, "d": 0.2}, 8, "run-time L and nb"),
    ("kagome", 4, "kagome-DM", {"j": 1.0, "d": 0.2}, 13, "stride 16, 4 points"),
    ("triangular", 7, "triangular-heisenberg", {"j": 1.0}, 20, "stride 20, 3 points"),
    ("cubic", 6, "cubic-heisenberg", {"j": 1.0}, 23, "stride 24, 2 points"),
    ("kagome", 6, "kagome-DM", {"j": 1.0, "d": 0.2}, 26, "stride 28, 2 points"),
    ("cubic", 7, "cubic-heisenberg", {"j": 1.0}, 31, "stride 32, 2 points"),
]


@pytest.mark.parametrize("case", SHAPE_CASES, ids=[c[5].replace(" ", "-").replace(",", "") for c in SHAPE_CASES])
def test_tri_compiled_shapes(case):
    """Every dGamma item and dSigma on a 4-point mesh against the oracle, at the
    initial state and at a frequency-dependent state (one oracle Euler step of
    it), for each bundle geometry of the 128-thread kernel."""
    from helpers import elem_rel, tri_problem_of
    from oracle import TriOracle
    from spinparser_amd.capi import TriFlowContext
    lattice, rng, model, cpl, L, shape = case
    p = tri_problem_of(lattice, rng, model, cpl, 4, 0.01, 20.0, 20.0, 0.05, 0.8)
    assert p.L == L
    o = TriOracle(p)
    ctx = TriFlowContext(p)
    v2, v4 = p.initial_state(np.float64)
    for k in (0, len(p.cutoffs) // 2):
        cut = float(p.cutoffs[k])
        if k > 0:
            dv2, dv4, _ = o.compute_step(cut, v2, v4, dtype=np.float64, nthreads=8)
            v2, v4 = v2 - 0.05 * cut * dv2, v4 - 0.05 * cut * dv4
        _log("TRI %s r%d L=%d O=%d (%s) at cutoff %.4g" % (lattice, rng, p.L, p.overlap_total, shape, cut))
        r2, r4, _ = o.compute_step(cut, v2, v4, dtype=np.float64, nthreads=8)
        ctx.set_state(cut, v2, v4)
        ctx.compute_step()
        g2, g4 = ctx.get_flow()
        e2, e4, el4 = max_rel(g2, r2), max_rel(g4, r4), elem_rel(g4, r4, 1e-3)
        _log("  dSigma %.3g  dGamma %.3g (elementwise %.3g)" % (e2, e4, el4))
        assert np.abs(r4).max() > 0
        assert e2 < TOL and e4 < TOL and el4 < TOL
    ctx.close()
