"""The C++ caller of the C ABI (tests/cpp/runcore_su2: FrgCore-shaped classes
whose flow object's isDiverged() calls sp_is_diverged and whose correlation
measurement reads the device state) runs SpinParser::runCore's loop
(src/SpinParser.cpp:141-172) on the hyperkagome golden problem.  Its records
must equal those of the Python FrgCore path bit for bit (same kernels, same
order), and the whole FP64 correlation flow must match the FP64 oracle's flow
to 1e-10: normwise per record and per channel, and elementwise relative to
each entry with a floor of 1e-6 of the record's largest entry (entries below
the floor are held to 1e-10 of it in absolute terms).
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from helpers import ref1_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "runcore_su2")
TOL = 1e-10
FLOOR = 1e-6


def _write_problem(prob, path):
    v2, v4s, v4d = prob.initial_state(np.float64)
    O = int(prob.ovl_off[-1])
    with open(path, "wb") as f:
        f.write(b"SPP1")
        f.write(struct.pack("<7i", prob.nw, prob.L, O, len(prob.range_fwd), prob.B, prob.Rb, len(prob.cutoffs)))
        f.write(np.ascontiguousarray(prob.mesh, dtype="<f4").tobytes())
        for a in (prob.site_rid, prob.inv_rid, prob.ovl_off, prob.ovl_rid1, prob.ovl_rid2, prob.range_fwd,
                  prob.range_inv, prob.corr_rid):
            f.write(np.ascontiguousarray(a, dtype="<i4").tobytes())
        f.write(struct.pack("<d", float(prob.spin)))
        for a in (np.asarray(prob.cutoffs, dtype=np.float64), v2, v4s, v4d):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def _read_records(path, n):
    with open(path, "rb") as f:
        nrec, div = struct.unpack("<2i", f.read(8))
        recs = []
        for _ in range(nrec):
            cut = struct.unpack("<d", f.read(8))[0]
            zz = np.frombuffer(f.read(8 * n), dtype="<f8")
            dd = np.frombuffer(f.read(8 * n), dtype="<f8")
            recs.append((cut, zz, dd))
    return recs, bool(div)


def _elementwise(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    floor = FLOOR * np.abs(b).max()
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor)).max()) if floor > 0 else float(np.abs(a - b).max())


@pytest.mark.gpu
def test_cpp_runcore_equals_python_and_oracle(tmp_path):
    from oracle import Oracle
    from spinparser_amd.frgcore import HipSU2FrgCore, SU2MeasurementCorrelation, run_core
    from spinparser_amd.task import MeasurementSpec
    assert os.path.exists(EXE), "tests/cpp/runcore_su2 is not built (run __graft_entry__.build())"
    prob = ref1_problem()
    print("C++ runCore on hyperkagome r3: L=%d nw=%d cutoffs=%d" % (prob.L, prob.nw, len(prob.cutoffs)), flush=True)
    sys.stdout.flush()
    pfile, ofile = str(tmp_path / "problem.bin"), str(tmp_path / "records.bin")
    _write_problem(prob, pfile)
    res = subprocess.run([EXE, pfile, ofile], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    n = prob.B * prob.Rb
    crecs, cdiv = _read_records(ofile, n)
    assert not cdiv
    assert len(crecs) == len(prob.cutoffs)

    m = SU2MeasurementCorrelation(MeasurementSpec("correlation"))
    core = HipSU2FrgCore(prob, [m])
    assert run_core(core, prob.cutoffs)
    core.close()
    assert len(m.records) == len(crecs)
    for (cut, zz, dd), r in zip(crecs, m.records):
        assert cut == r.cutoff
        assert np.array_equal(zz, r.zz.ravel()) and np.array_equal(dd, r.dd.ravel())

    # the whole FP64 flow against the FP64 oracle flow
    ozz, odd, on, _ = Oracle(prob).run_flow(np.float64)
    assert on == len(crecs)
    worst_norm, worst_elem = 0.0, 0.0
    for k, (_, zz, dd) in enumerate(crecs):
        for got, ref in ((zz, ozz[k].ravel()), (dd, odd[k].ravel())):
            den = np.abs(ref).max()
            if den == 0.0:  # the first record at cutoff = max frequency vanishes identically
                assert np.abs(got).max() == 0.0
                continue
            worst_norm = max(worst_norm, float(np.abs(got - ref).max() / den))
            worst_elem = max(worst_elem, _elementwise(got, ref))
    print("SU2 whole FP64 flow vs oracle: normwise %.3g, elementwise (floor %.0e) %.3g over %d records"
          % (worst_norm, FLOOR, worst_elem, len(crecs)), flush=True)
    assert worst_norm < TOL
    assert worst_elem < TOL


def test_cpp_runcore_reports_bad_problem(tmp_path):
    """A malformed problem file is an error with exit code 1, not a crash (CPU:
    the file is rejected before any HIP call)."""
    assert os.path.exists(EXE), "tests/cpp/runcore_su2 is not built (run __graft_entry__.build())"
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"SPP1" + struct.pack("<7i", 4, 2, 0, 1, 1, 1, 3))
    res = subprocess.run([EXE, str(bad), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert res.returncode == 1
    assert "truncated" in res.stderr
