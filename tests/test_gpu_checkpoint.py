"""Checkpoint resume on the GPU cores (test/scripted/test_checkpoint.sh:101-105
in spirit): a flow interrupted by writeCheckpoint / readCheckpoint into a new
core reproduces the uninterrupted flow's measurements.  The checkpoint stores
the state as float, like the reference, so the resumed FP64 flow agrees to
float rounding (the reference's own tolerance, 1e-5 absolute, is the bar).
"""
import numpy as np
import pytest

from helpers import ref1_problem, tri_ref3_problem, xyz_ref2_problem

pytestmark = pytest.mark.gpu


def _flow(core, cutoffs, start, stop):
    recs = []
    for k in range(start, stop):
        core.computeStep()
        recs.append(core.context.measure_correlation())
        assert not core.flow().isDiverged()
        core.finalizeStep(float(cutoffs[k + 1]))
    return recs


def _flat(r):
    return np.concatenate([np.ravel(x) for x in r]) if isinstance(r, tuple) else np.ravel(r)


@pytest.mark.parametrize("core_kind", ["SU2", "TRI", "XYZ"])
def test_resume_from_checkpoint(tmp_path, core_kind):
    from spinparser_amd.frgcore import HipSU2FrgCore, HipTRIFrgCore, HipXYZFrgCore
    p = {"SU2": ref1_problem, "TRI": tri_ref3_problem, "XYZ": xyz_ref2_problem}[core_kind]()
    cls = {"SU2": HipSU2FrgCore, "TRI": HipTRIFrgCore, "XYZ": HipXYZFrgCore}[core_kind]
    make = lambda: cls(p, [])  # noqa: E731
    path = str(tmp_path / "flow.data")
    n = len(p.cutoffs) - 1
    mid = n // 2
    core = make()
    _flow(core, p.cutoffs, 0, mid)
    assert core.writeCheckpoint(path) == 0
    full = _flow(core, p.cutoffs, mid, n)
    core.close()
    resumed = make()
    assert resumed.readCheckpoint(path)
    assert resumed.cutoff == pytest.approx(float(p.cutoffs[mid]))
    again = _flow(resumed, p.cutoffs, mid, n)
    resumed.close()
    worst = max(float(np.abs(_flat(a) - _flat(b)).max()) for a, b in zip(again, full))
    print("%s resume: max abs difference %.3g over %d records" % (core_kind, worst, len(full)), flush=True)
    assert worst < 1e-5
