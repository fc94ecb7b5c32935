"""Host-side checks (no GPU): lattice factory known answers, task grids, work
counter, and that the C-ABI library loads and exports every declared symbol."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import cfg1_problem, cfg2_problem, ref1_problem
from spinparser_amd.lattice import (LatticeUnitCell, SpinModelUnitCell, new_lattice_model, string_to_double)
from spinparser_amd.task import exponential_cutoffs, exponential_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _square_r3():
    """Lattice of test/test_Lattice.cpp:9-44 (square, range 3, Heisenberg)."""
    uc = LatticeUnitCell(lattice_vectors=[(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)],
                         basis_sites=[(0.0, 0.0, 0.0)], bonds=[(0, 0, 1, 0, 0), (0, 0, 0, 1, 0)])
    m = SpinModelUnitCell()
    m.add((0, 0, 0, 0), (1, 0, 0, 0), np.eye(3, dtype=np.float32))
    m.add((0, 0, 0, 0), (0, 1, 0, 0), np.eye(3, dtype=np.float32))
    return new_lattice_model(uc, m, 3)


def _honeycomb_r3():
    """Lattice of test/test_Lattice.cpp:46-84 (honeycomb, anisotropic diagonal couplings)."""
    uc = LatticeUnitCell(lattice_vectors=[(1.5, 0.5 * np.sqrt(3), 0.0), (1.5, -0.5 * np.sqrt(3), 0.0), (0.0, 0.0, 1.0)],
                         basis_sites=[(0.0, 0.0, 0.0), (1.0, 0.0, 0.0)],
                         bonds=[(0, 1, 0, 0, 0), (1, 0, 1, 0, 0), (1, 0, 0, 1, 0)])
    m = SpinModelUnitCell()
    for to, (a, v) in (((0, 0, 0, 1), (0, 0.1)), ((0, -1, 0, 1), (1, 0.2)), ((-1, 0, 0, 1), (2, 0.2))):
        st = np.zeros((3, 3), dtype=np.float32)
        st[a, a] = np.float32(v)
        m.interactions.append(__import__("spinparser_amd.lattice", fromlist=["SpinInteraction"]).SpinInteraction(
            (0, 0, 0, 0), to, st))
    return new_lattice_model(uc, m, 3)


def _idx(lat, p):
    return lat.sites.index(p)


def test_square_lattice_sizes():
    lat = _square_r3()
    assert len(lat.ranges[0]) == 25 and lat.size == 6
    assert lat.sites[lat.basis_ids[0]] == (0, 0, 0, 0) and lat.sites[0] == (0, 0, 0, 0)


def test_square_lattice_symmetry_table():
    """A sample of the pairs of test_Lattice.cpp:117-124: sym(i1,i2) == sym(0,target)."""
    lat = _square_r3()
    cases = [((0, 0), (0, -1), (0, -1)), ((0, 0), (-1, 0), (0, -1)), ((0, 0), (1, 1), (-1, -1)),
             ((0, 0), (2, 1), (-1, -2)), ((0, 0), (0, 3), (0, -3)), ((1, 0), (0, 0), (0, -1)),
             ((1, 0), (-1, 0), (0, -2)), ((1, 0), (1, 0), (0, 0)), ((1, 0), (1, -2), (0, -2))]
    for a, b, t in cases:
        i1, i2, tg = (_idx(lat, (x, y, 0, 0)) for x, y in (a, b, t))
        assert lat.symmetry_transform(i1, i2) == lat.symmetry_transform(0, tg)


def test_square_lattice_overlap():
    """getOverlap(rid(1,1)) has 18 terms forming the multiset of test_Lattice.cpp:130-131."""
    lat = _square_r3()
    i1 = [(0, 0), (0, -1), (0, -2), (-1, -1), (0, -3), (-1, -2), (0, -1), (-1, -1), (-1, -1), (0, -2), (0, -1),
          (0, -1), (-1, -2), (-1, -2), (-1, -2), (-1, -2), (-1, -2), (0, -3)]
    i2 = [(-1, -1), (0, -1), (-1, -1), (0, 0), (-1, -2), (0, -1), (0, -1), (0, -2), (0, -2), (-1, -1), (-1, -2),
          (-1, -2), (0, -3), (-1, -2), (-1, -2), (0, -1), (0, -3), (-1, -2)]
    want = sorted((lat.symmetry_transform(0, _idx(lat, (a[0], a[1], 0, 0))),
                   lat.symmetry_transform(0, _idx(lat, (b[0], b[1], 0, 0)))) for a, b in zip(i1, i2))
    r11 = lat.symmetry_transform(0, _idx(lat, (1, 1, 0, 0)))
    ov = lat.overlaps[r11]
    assert len(ov[0]) == 18
    assert sorted(zip(ov[0].tolist(), ov[1].tolist())) == want


_HC_I1 = [(0, 0, 0)] * 19 + [(1, 0, 0)] * 12
_HC_I2 = [(0, 0, 0), (0, 0, 1), (-1, 0, 1), (1, 0, 0), (-1, 0, 0), (-1, 1, 0), (1, 0, 1), (1, -1, 1), (-2, 0, 1),
          (-1, -1, 1), (-2, 1, 1), (0, 1, 0), (0, 1, 1), (-1, 1, 1), (0, -1, 0), (1, -1, 0), (0, -1, 1), (0, -2, 1),
          (1, -2, 1), (0, 0, 0), (0, 0, 1), (-1, 0, 1), (1, 0, 0), (1, 0, 1), (1, -1, 1), (0, 1, 0), (0, 1, 1),
          (-1, 1, 1), (1, -1, 0), (0, -1, 1), (1, -2, 1)]
_HC_T = [(0, 0, 0), (0, 0, 1), (-1, 0, 1), (1, 0, 0), (-1, 0, 0), (-1, 1, 0), (1, 0, 1), (1, -1, 1), (-2, 0, 1),
         (-1, -1, 1), (-2, 1, 1), (1, 0, 0), (1, 0, 1), (1, -1, 1), (-1, 0, 0), (-1, 1, 0), (-1, 0, 1), (-2, 0, 1),
         (-2, 1, 1), (-1, 0, 0), (-1, 0, 1), (-2, 0, 1), (0, 0, 0), (0, 0, 1), (-1, 0, 1), (-1, 1, 0), (1, -1, 1),
         (-2, 1, 1), (-1, 0, 0), (-1, -1, 1), (-2, 0, 1)]
_HC_SY = "YYYYYYYYYYYZZZZZZZZYYYYYZYZYZYZ"
_HC_SZ = "ZZZZZZZZZZZYYYYYYYYZZZZZYZYZYZY"


def test_honeycomb_lattice_with_spin_permutations():
    """test_Lattice.cpp:190-240: 19 sites in range, 11 representatives, and the
    31 symmetry maps with their spin permutations: symmetryTransform(i1, i2,
    x, y, z) == symmetryTransform(0, target, x, sy, sz)."""
    lat = _honeycomb_r3()
    assert len(lat.ranges[0]) == 19 and lat.size == 11
    comp = {"X": 0, "Y": 1, "Z": 2}
    for n in range(31):
        a, b, t = _HC_I1[n], _HC_I2[n], _HC_T[n]
        i1 = _idx(lat, (a[0], a[1], 0, a[2]))
        i2 = _idx(lat, (b[0], b[1], 0, b[2]))
        tg = _idx(lat, (t[0], t[1], 0, t[2]))
        assert lat.symmetry_transform(i1, i2) == lat.symmetry_transform(0, tg), n
        p = lat.sym_perm[i1, i2].tolist()
        q = lat.sym_perm[0, tg].tolist()
        assert [p[0], p[1], p[2]] == [q[0], q[comp[_HC_SY[n]]], q[comp[_HC_SZ[n]]]], n


def test_survey_sizes():
    """Representative / range / overlap counts of the benchmark lattices."""
    p1 = cfg1_problem(nw=4)
    assert (p1.L, len(p1.range_fwd), p1.overlap_total) == (9, 41, 226)
    p2 = cfg2_problem(nw=4)
    assert (p2.L, len(p2.range_fwd), p2.overlap_total) == (23, 377, 4095)


def test_expression_parser():
    assert string_to_double("sqrt(3)/2") == pytest.approx(np.sqrt(3) / 2, rel=1e-15)
    assert string_to_double("7/2/sqrt(2)") == pytest.approx(3.5 / np.sqrt(2), rel=1e-15)
    assert string_to_double("-1/sqrt(2)") == pytest.approx(-1 / np.sqrt(2), rel=1e-15)
    assert string_to_double("2*sqrt(2)") == pytest.approx(2 * np.sqrt(2), rel=1e-15)


def test_grids():
    m = exponential_mesh(0.005, 50.0, 32)
    assert m.dtype == np.float32 and len(m) == 32 and m[0] == np.float32(0.005)
    assert abs(m[-1] - 50.0) < 1e-4 and np.all(np.diff(m) > 0)
    c = exponential_cutoffs(50.0, 0.01, 0.98)
    assert c.dtype == np.float32 and c[0] == np.float32(50.0) and c[-1] > np.float32(0.01)
    assert len(c) == 422
    assert len(exponential_cutoffs(100.0, 0.005, 0.985)) == 656
    # golden cutoffs of test_reference1 are reproduced bit for bit (checked in test_oracle_golden)
    assert ref1_problem().cutoffs[1] == np.float32(np.float32(10.0) * np.float32(0.9))


def test_work_counter_matches_oracle():
    from oracle import Oracle
    from spinparser_amd.workload import evaluation_calls
    prob = cfg1_problem(nw=6)
    o = Oracle(prob)
    st = prob.initial_state(np.float64)
    for k in (0, 50, 200, 420):
        _, _, _, calls = o.compute_step(prob.cutoffs[k], *st)
        assert tuple(calls.tolist()) == evaluation_calls(prob.mesh, prob.cutoffs[k])


def test_library_exports_header_symbols():
    """libspfrg.so loads without a GPU and exports every function of include/spfrg.h."""
    from spinparser_amd.build import LIB_PATH
    from spinparser_amd.capi import EXPORTED_SYMBOLS, TRI_EXPORTED_SYMBOLS, XYZ_EXPORTED_SYMBOLS
    with open(os.path.join(ROOT, "include", "spfrg.h")) as fh:
        header = fh.read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(sp_\w+)\s*\(", header, flags=re.M))
    assert declared == set(EXPORTED_SYMBOLS) | set(TRI_EXPORTED_SYMBOLS) | set(XYZ_EXPORTED_SYMBOLS)
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_invalid_problem_rejected_without_gpu_work():
    """sp_ctx_footprint validates its descriptor on the host."""
    from spinparser_amd.capi import FlowError, footprint_bytes
    prob = cfg1_problem(nw=8)
    assert footprint_bytes(prob) > 2 * prob.vertex_size * 16


def test_bench_stratified_sampling_and_estimator():
    """The bench line's sampled rate: strata midpoints over the schedule, and
    the whole-schedule time from seconds per algorithmic FLOP interpolated
    between the samples (exact when that rate is constant or linear in the
    evaluation index)."""
    import bench
    n = 655
    ks = bench.sample_indices(n, 20)
    assert len(ks) == 20 and ks == sorted(set(ks)) and ks[0] >= 0 and ks[-1] < n
    assert ks[0] < n / 20 and ks[-1] >= n - n / 20
    assert bench.sample_indices(n, n) == list(range(n))
    rng = np.random.default_rng(1)
    flops = rng.uniform(1e9, 4e11, n)
    spf = 2.5e-14
    exact = spf * flops.sum()
    assert abs(bench.full_schedule_seconds(flops, ks, [spf * flops[k] for k in ks]) - exact) <= 1e-12 * exact
    a, b = 1e-14, 3e-17
    t = [(a + b * k) * flops[k] for k in ks]
    want = sum(flops[i] * (a + b * min(max(i, ks[0]), ks[-1])) for i in range(n))
    assert abs(bench.full_schedule_seconds(flops, ks, t) - want) <= 1e-9 * want


def test_bench_traffic_constants_match_profiles():
    """The roofline traffic of the bench line comes from the FETCH_SIZE and
    WRITE_SIZE constants in bench.py; each must be the figure that the
    committed rocprofv3 summary of the same kernel and workload reports (the
    first row of its HBM traffic table, KiB per dispatch, rounded there)."""
    import re
    import bench
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    files = {("cubic-j1j2-r6-nw64", "f64"): "k_dgamma_cubic-j1j2-r6-nw64_fp64.md",
             ("square-r4-nw32", "f64"): "k_dgamma_square-r4-nw32_fp64.md",
             ("kagome-dm-r5-nw48", "f64"): "k_tri_dgamma_kagome-dm-r5-nw48_fp64.md",
             ("honeycomb-kitaev-r8-nw64", "f64"): "k_xyz_dgamma_honeycomb-kitaev-r8-nw64_fp64.md"}
    assert set(files) == set(bench.TRAFFIC_KIB)
    for key, name in files.items():
        text = open(os.path.join(root, "profiles", name)).read()
        head = text.index("| kernel | FETCH_SIZE | WRITE_SIZE |")
        row = text[head:].split("\n")[2]
        fetch, write = (float(x.replace(",", "")) for x in re.findall(r"\| ([0-9][0-9,.]*) ", row + " ")[:2])
        t = bench.TRAFFIC_KIB[key]
        assert abs(t["fetch_kib"] - fetch) <= max(1.0, 1e-4 * fetch), (name, t["fetch_kib"], fetch)
        assert abs(t["write_kib"] - write) <= max(1.0, 1e-4 * write), (name, t["write_kib"], write)


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus N > 1 exits with a message before any GPU work."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_task_status_round_trip_and_corruption(tmp_path):
    """The <calculation> element of the task file (src/TaskFileParser.cpp:33-72,
    349-374): written and read back for every status; a status without the
    times it requires is rejected; 'new' leaves no element behind."""
    from spinparser_amd.frgcore import InitializationError
    from spinparser_amd.runner import FINISHED, NEW, POSTPROCESSING, RUNNING, ComputationStatus, read_status, \
        write_task_file
    from spinparser_amd.task import make_task_xml
    path = str(tmp_path / "task.xml")
    with open(path, "w") as f:
        f.write(make_task_xml(lattice="square", lattice_range=2, model="square-heisenberg", symmetry="SU2",
                              couplings={"j": 1.0}, freq_min=0.1, freq_max=10.0, freq_count=4, cutoff_max=10.0,
                              cutoff_min=1.0, cutoff_step=0.9))
    assert read_status(path).identifier == NEW
    for st in (RUNNING, POSTPROCESSING, FINISHED):
        want = ComputationStatus(st, "2024-Jan-05 13:02:11", "2024-Jan-05 13:05:00",
                                 "2024-Jan-05 13:06:00" if st == FINISHED else None)
        write_task_file(path, want)
        got = read_status(path)
        assert got == want
    write_task_file(path, ComputationStatus(NEW))
    assert read_status(path).identifier == NEW
    with open(path) as f:
        assert "<calculation" not in f.read()
    # a finished task without an end time, a running one without a start time
    import xml.etree.ElementTree as ET
    for attrs in ({"status": FINISHED, "startTime": "a", "checkpointTime": "b"},
                  {"status": RUNNING}, {"status": "bogus", "startTime": "a"}):
        tree = ET.parse(path)
        root = tree.getroot()
        for old in root.findall("calculation"):
            root.remove(old)
        ET.SubElement(root, "calculation", attrs)
        tree.write(path)
        with pytest.raises(InitializationError):
            read_status(path)
