"""TRI oracle pins (CPU).

* the reference's TRI golden flow (test/scripted/assets/test_reference3.ref:
  honeycomb Heisenberg-Kitaev-Gamma, range 3, 14 frequencies, 34 cutoffs, all 10
  TRICor* groups; the reference's own tolerance is 1e-5 absolute,
  test/scripted/assets/test_eval.py:14,28-29) -- float instantiation;
* TRI on a Heisenberg model equals the SU(2) core (test/scripted/test_reference1.sh:71-78):
  XX = YY = ZZ = SU2CorZZ and DD = SU2CorDD;
* structure of the Pauli-derived coefficient tables: term counts, and the
  density-density particle-particle row, whose form is fixed by sigma^a sigma^a = 1;
* antisymmetric (Dzyaloshinskii-Moriya) couplings, which no reference fixture
  holds, through the spin symmetries they must respect: Heisenberg plus an
  out-of-plane DM term is invariant under spin rotations about z, so every
  correlation obeys XX = YY, XY = -YX and XZ = YZ = ZX = ZY = 0; and the global
  pi rotation about x maps the model at D onto the model at -D, so XX, YY, ZZ,
  DD, YZ, ZY are even in D and XY, YX, XZ, ZX odd.  A wrong sign, phase or
  component index in the off-diagonal coupling algebra breaks these relations.
"""
import numpy as np
import pytest

from helpers import GOLDEN, manual_task_xml, tri_ref3_problem
from oracle import TRI_GROUPS, Oracle, TriOracle


@pytest.fixture(scope="module")
def ref3_flow():
    p = tri_ref3_problem()
    rec, n, _ = TriOracle(p).run_flow(dtype=np.float32, nthreads=8)
    return p, rec, n


def test_tri_golden_flow(ref3_flow):
    p, rec, n = ref3_flow
    g = np.load(GOLDEN + "/reference3_tri.npz")
    assert n == len(p.cutoffs) == 34
    worst = 0.0
    for k in range(n):
        for gi, grp in enumerate(TRI_GROUPS):
            assert abs(g["%s/data/measurement_%d/cutoff" % (grp, k)][0] - p.cutoffs[k]) == 0.0
            ref = g["%s/data/measurement_%d/data" % (grp, k)]
            worst = max(worst, float(np.abs(rec[k, gi] - ref).max()))
    # float arithmetic of a different operation order: stays at rounding level
    assert worst < 1e-6, worst


def test_tri_golden_site_order(ref3_flow):
    p, _, _ = ref3_flow
    g = np.load(GOLDEN + "/reference3_tri.npz")
    sites = g["TRICorXX/meta/sites"]
    lat = p.lattice
    for b in range(p.B):
        for k, j in enumerate(lat.ranges[b]):
            assert np.allclose(lat.position(j), sites[b, k], atol=1e-5)


def test_tri_heisenberg_equals_su2():
    from spinparser_amd.problem import su2_problem, tri_problem
    from spinparser_amd.task import parse_task
    from helpers import GOLDEN_RES
    xs = manual_task_xml("honeycomb", 2, "honeycomb-heisenberg", {"j": "1.0"}, cmin="1.0")
    xt = manual_task_xml("honeycomb", 2, "honeycomb-heisenberg", {"j": "1.0"}, symmetry="TRI", cmin="1.0")
    ps = su2_problem(parse_task(xs, True), GOLDEN_RES)
    pt = tri_problem(parse_task(xt, True), GOLDEN_RES)
    zz, dd, n, _ = Oracle(ps).run_flow(dtype=np.float64, nthreads=8)
    rec, nt, _ = TriOracle(pt).run_flow(dtype=np.float64, nthreads=8)
    assert n == nt and n > 5
    scale = np.abs(zz[:n]).max()
    for gi, grp in enumerate(TRI_GROUPS):
        if grp == "TRICorDD":
            assert np.abs(rec[:n, gi] - dd[:n]).max() < 1e-12 * max(1.0, np.abs(dd[:n]).max())
        elif grp[-1] == grp[-2]:
            assert np.abs(rec[:n, gi] - zz[:n]).max() < 1e-12 * scale, grp
        else:
            assert np.abs(rec[:n, gi]).max() < 1e-12 * scale, grp


def test_tri_tables_structure():
    counts = [TriOracle.table(k)[1] for k in range(5)]
    assert counts == [256, 256, 256, 256, 64]
    S, _ = TriOracle.table(0)
    # particle-particle density-density row: sum over components c of
    # zeta-weighted products A[c] B[c] (sigma^a sigma^a = 1 on both lines; the
    # i of mixed components squares to -1)
    row = S[15]
    for a in range(16):
        for b in range(16):
            if a != b:
                assert row[a, b] == 0.0
            else:
                mu, nu = divmod(a, 4)
                expect = (-1.0 if (mu == 3) != (nu == 3) else 1.0)
                assert row[a, a] == expect
    # every coefficient is +-1 (+-2 in the lattice bubble)
    for k in range(4):
        t, _ = TriOracle.table(k)
        assert set(np.unique(np.abs(t))) <= {0.0, 1.0}
    t, _ = TriOracle.table(4)
    assert set(np.unique(np.abs(t))) == {0.0, 2.0}


def test_tri_dm_spin_symmetries():
    """Kagome J = 1 with out-of-plane DM (model kagome-DM of spinparser_amd/res), range 3."""
    from helpers import tri_problem_of
    flows = {}
    for d in (0.3, -0.3):
        p = tri_problem_of("kagome", 3, "kagome-DM", {"j": 1.0, "d": d}, 8, 0.05, 20.0, 10.0, 0.8, 0.9)
        rec, n, _ = TriOracle(p).run_flow(dtype=np.float64, nthreads=8)
        assert n == len(p.cutoffs) - 1 or n == len(p.cutoffs)
        flows[d] = rec[:n]
    g = {name: i for i, name in enumerate(TRI_GROUPS)}
    a, b = flows[0.3], flows[-0.3]
    assert a.shape == b.shape
    scale = np.abs(a).max()
    tol = 1e-12 * scale
    # the DM term is felt: the off-diagonal response is non-zero and SU(2) is broken
    assert np.abs(a[:, g["TRICorXY"]]).max() > 1e-3 * scale
    assert np.abs(a[:, g["TRICorXX"]] - a[:, g["TRICorZZ"]]).max() > 1e-4 * scale
    # U(1) about z
    assert np.abs(a[:, g["TRICorXX"]] - a[:, g["TRICorYY"]]).max() < tol
    assert np.abs(a[:, g["TRICorXY"]] + a[:, g["TRICorYX"]]).max() < tol
    for name in ("TRICorXZ", "TRICorYZ", "TRICorZX", "TRICorZY"):
        assert np.abs(a[:, g[name]]).max() < tol, name
    # pi rotation about x: D -> -D
    parity = {"TRICorDD": 1, "TRICorXX": 1, "TRICorXY": -1, "TRICorXZ": -1, "TRICorYX": -1,
              "TRICorYY": 1, "TRICorYZ": 1, "TRICorZX": -1, "TRICorZY": 1, "TRICorZZ": 1}
    for name, sgn in parity.items():
        assert np.abs(a[:, g[name]] - sgn * b[:, g[name]]).max() < tol, name


def test_kagome_dm_initial_orientation():
    """The DM term's orientation in the initial TRI vertex, derived by hand from
    the reference's rules.  kagome-DM (res/models.xml:145-170) writes each bond
    twice, "xy" with +d and "-yx" with -d (src/LatticeModelFactory.cpp:321-336:
    strength[s1][s2] += sign * d).  At range 1 the one representative
    neighbour of the origin (0,0,0,0) is (-1,0,0,1); the model's bond
    from="-1,0,0,1" to="0,0,0,0" connects it *to* the origin, so
    isConnectingSites(origin, site) = -1 (src/LatticeModelFactory.cpp:196-207)
    and the coupling enters transposed (:976-982): J(0 -> rid)[x][y] = -d,
    J[y][x] = +d.  TRIEffectiveAction sets 0.25 J[s1][s2] / normalization at
    every frequency (src/TRI/TRIEffectiveAction.hpp:38-64; normalization 1), so
    with j = 1, d = 0.2: xx = yy = zz = 0.25, xy = -0.05, yx = +0.05, all other
    components 0, and the on-site vertex (rid 0) is zero."""
    from helpers import tri_problem_of
    p = tri_problem_of("kagome", 1, "kagome-DM", {"j": 1.0, "d": 0.2}, 4, 0.05, 10.0, 10.0, 0.1, 0.9)
    assert p.L == 2
    assert tuple(p.lattice.sites[1]) == (-1, 0, 0, 1)
    _, v4 = p.initial_state(np.float64)
    L = p.L
    g = v4.reshape(p.n_freq, 16, L)
    expect = np.zeros((16, L))
    x, y, z = 0, 1, 2
    for s1, s2, val in ((x, x, 0.25), (y, y, 0.25), (z, z, 0.25), (x, y, -0.05), (y, x, 0.05)):
        expect[4 * s1 + s2, 1] = val
    assert np.allclose(g, expect[None, :, :], rtol=0, atol=1e-7)
    # the float instantiation carries the same orientation
    _, v4f = p.initial_state(np.float32)
    assert np.allclose(v4f.reshape(p.n_freq, 16, L), expect[None, :, :], rtol=0, atol=1e-7)
