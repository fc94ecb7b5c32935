"""Pins the CPU oracle against the reference's own golden vectors and known answers.

* test/scripted/assets/test_reference1.ref (SU2 hyperkagome flow, float32)
* test/scripted/assets/test_python_obs.py literals (SU2 honeycomb flow)
* test/test_Integrator.cpp known answers
* the reference's deliberately failing comparison (test_referenceFail.sh, J=-1)
* closed-form special cases: Lambda >= w_max gives zero susceptibility; the
  initial state without on-site coupling has dSigma = 0.
"""
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, cfg1_problem, max_rel, python_obs_problem, ref1_problem
from oracle import Oracle


@pytest.fixture(scope="module")
def golden1():
    return np.load(os.path.join(GOLDEN, "reference1_su2.npz"))


@pytest.fixture(scope="module")
def ref1_flow():
    prob = ref1_problem()
    zz, dd, n, _ = Oracle(prob).run_flow(np.float32)
    return prob, zz, dd, n


def test_ref1_site_order_matches_golden(golden1):
    prob = ref1_problem()
    lat = prob.lattice
    sites = np.array([[lat.position(j) for j in lat.ranges[b]] for b in range(prob.B)], dtype=np.float32)
    np.testing.assert_array_equal(sites, golden1["SU2CorZZ/meta/sites"])
    basis = np.array([lat.position(i) for i in lat.basis_ids], dtype=np.float32)
    np.testing.assert_array_equal(basis, golden1["SU2CorZZ/meta/basis"])


def test_ref1_flow_matches_golden(golden1, ref1_flow):
    prob, zz, dd, n = ref1_flow
    n_ref = len([k for k in golden1.files if k.startswith("SU2CorZZ/data/") and k.endswith("/data")])
    assert n == n_ref == len(prob.cutoffs)
    for k in range(n):
        assert np.float32(golden1["SU2CorZZ/data/measurement_%d/cutoff" % k][0]) == prob.cutoffs[k]
        # the reference comparator's tolerance (test/scripted/assets/test_eval.py:14)
        assert np.abs(zz[k] - golden1["SU2CorZZ/data/measurement_%d/data" % k]).max() < 1e-5
        assert np.abs(dd[k] - golden1["SU2CorDD/data/measurement_%d/data" % k]).max() < 1e-5


def test_ref1_flow_is_bit_exact(golden1, ref1_flow):
    """The float instantiation reproduces the reference arithmetic exactly."""
    _, zz, dd, n = ref1_flow
    for k in range(n):
        np.testing.assert_array_equal(zz[k], golden1["SU2CorZZ/data/measurement_%d/data" % k])
        np.testing.assert_array_equal(dd[k], golden1["SU2CorDD/data/measurement_%d/data" % k])


def test_reference_fail_case_is_detected(golden1):
    """test_referenceFail.sh: the same flow with J=-1 must not match the golden file."""
    prob = ref1_problem(j="-1.0")
    zz, dd, n, _ = Oracle(prob).run_flow(np.float32)
    worst = max(np.abs(zz[k] - golden1["SU2CorZZ/data/measurement_%d/data" % k]).max() for k in range(n))
    assert worst > 1e-5


def test_python_obs_literals():
    """Honeycomb SU2 flow values held as literals in test_python_obs.py."""
    with open(os.path.join(GOLDEN, "python_obs_su2.json")) as fh:
        lit = json.load(fh)
    prob = python_obs_problem()
    zz, dd, n, _ = Oracle(prob).run_flow(np.float32)
    lat = prob.lattice
    pos0 = np.array([lat.position(j) for j in lat.ranges[0]])
    np.testing.assert_allclose(pos0, np.array(lit["sites_ref0"]), atol=1e-6)
    k100 = int(np.argmin(np.abs(pos0 - np.array([1.0, 0.0, 0.0])).sum(axis=1)))
    # XX component of the SU2 correlation is the ZZ record (obs.py component mapping)
    np.testing.assert_allclose(zz[:, 0, k100], lit["corr_site100_xx"], rtol=1e-5, atol=0)
    # site-resolved ZZ at the last cutoff (0.309031)
    np.testing.assert_allclose(zz[n - 1, 0, :], lit["corr_zz_cutoff0309031"], rtol=1e-5, atol=0)


@pytest.mark.parametrize("kind,a,b,lin,quad", [
    (0, 1.5, 0.0, 48.875, 333.5625),   # obscure left boundary [1.5, 10]
    (1, 0.0, 9.5, 44.625, 286.8125),   # obscure right boundary [1, 9.5]
    (2, 1.5, 9.5, 44.0, 285.875),      # obscure boundaries [1.5, 9.5]
])
def test_integrator_known_answers(kind, a, b, lin, quad):
    """test/test_Integrator.cpp:39-119 on the mesh 1..10."""
    out = Oracle.integrate_test(np.arange(1, 11, dtype=np.float32), kind, a, b)
    assert out[0] == pytest.approx(lin, abs=1e-6 * lin)
    assert out[1] == pytest.approx(quad, abs=1e-6 * quad)


def test_first_measurement_vanishes_above_mesh():
    """At Lambda = w_max every integration window is empty: chi = 0 (measurement_0)."""
    prob = ref1_problem()
    o = Oracle(prob)
    v2, v4s, v4d = prob.initial_state(np.float64)
    zz, dd = o.correlation(prob.cutoffs[0], v2, v4s, v4d)
    assert prob.cutoffs[0] == prob.mesh[-1]
    assert np.all(zz == 0.0) and np.all(dd == 0.0)


def test_initial_dsigma_vanishes():
    """Gamma_d = 0, Sigma = 0 and no on-site coupling: dSigma is exactly zero."""
    prob = cfg1_problem(nw=8)
    o = Oracle(prob)
    v2, v4s, v4d = prob.initial_state(np.float64)
    dv2, dss, ddd, _ = o.compute_step(prob.cutoffs[10], v2, v4s, v4d)
    assert np.all(dv2 == 0.0)
    assert np.abs(dss).max() > 0.0


def test_fp64_close_to_fp32():
    """Both instantiations of the oracle agree to single precision."""
    prob = cfg1_problem(nw=8)
    o = Oracle(prob)
    v2, v4s, v4d = prob.initial_state(np.float64)
    cut = prob.cutoffs[60]
    d64 = o.compute_step(cut, v2, v4s, v4d, np.float64)
    d32 = o.compute_step(cut, v2, v4s, v4d, np.float32)
    assert max_rel(d32[1], d64[1]) < 1e-5
    assert max_rel(d32[2], d64[2]) < 1e-5
