"""Deferred measurements through the task lifecycle (test/scripted/test_defer.sh
restated): for each core, the same task is run once with the correlation
measured inline and once with method="defer".  The deferred task ends its
first run in the post-processing status with the state of every measured
cutoff in its data file; the second run replays them through the measurement
and finishes.  Both observable files must agree to the reference's tolerance,
1e-5 absolute (test/scripted/assets/test_eval.py:14), with the same number of
measurements and the same cutoffs.  The data file stores the state as float,
as the reference does, so the deferred values differ from the inline ones by
float rounding of the state.
"""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from helpers import REF1_FREQS

pytestmark = pytest.mark.gpu

GROUPS = {"SU2": ("SU2CorZZ", "SU2CorDD"),
          "XYZ": ("XYZCorXX", "XYZCorYY", "XYZCorZZ", "XYZCorDD"),
          "TRI": ("TRICorDD", "TRICorXX", "TRICorXY", "TRICorXZ", "TRICorYX", "TRICorYY", "TRICorYZ",
                  "TRICorZX", "TRICorZY", "TRICorZZ")}


def _task(core, method):
    vals = "".join("<value>%s</value>" % v for v in REF1_FREQS.split())
    return ("<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<task><parameters>"
            "<frequency discretization=\"manual\">%s</frequency>"
            "<cutoff discretization=\"exponential\"><max>10</max><min>0.3</min><step>0.9</step></cutoff>"
            "<lattice name=\"triangular\" range=\"3\"/>"
            "<model name=\"triangular-heisenberg\" symmetry=\"%s\"><j>1.0</j></model>"
            "</parameters><measurements><measurement name=\"correlation\" method=\"%s\" /></measurements></task>"
            % (vals, core, method))


def _status(path):
    calc = ET.parse(path).getroot().find("calculation")
    return None if calc is None else calc.get("status")


@pytest.mark.parametrize("core", ["SU2", "XYZ", "TRI"])
def test_defer_equals_inline(tmp_path, core):
    from spinparser_amd.obs import checkpoint_count, read_correlations
    from spinparser_amd.runner import run_task
    from spinparser_amd.task import parse_task
    print("defer case %s" % core, flush=True)
    sys.stdout.flush()
    nd = str(tmp_path / ("test_defer.%s.NDEFER.xml" % core))
    df = str(tmp_path / ("test_defer.%s.DEFER.xml" % core))
    with open(nd, "w") as f:
        f.write(_task(core, ""))
    with open(df, "w") as f:
        f.write(_task(core, "defer"))
    assert run_task(nd) == "finished"
    assert _status(nd) == "finished"
    assert not os.path.exists(nd[:-4] + ".data")
    assert run_task(df) == "postprocessing"
    assert _status(df) == "postprocessing"
    ncut = len(parse_task(df).cutoffs)
    assert checkpoint_count(df[:-4] + ".data") == ncut  # one state per measured cutoff
    assert not os.path.exists(df[:-4] + ".obs")        # nothing measured inline
    assert run_task(df) == "finished"
    assert _status(df) == "finished"
    assert run_task(df) == "finished"                  # nothing left to do
    from spinparser_amd.frgcore import new_core_from_task
    c = new_core_from_task(parse_task(nd))
    shape = (c.problem.B, c.problem.Rb)
    c.close()
    worst = 0.0
    for g in GROUPS[core]:
        cut_a, a = read_correlations(nd[:-4] + ".obs", g, shape)
        cut_b, b = read_correlations(df[:-4] + ".obs", g, shape)
        assert len(cut_a) == ncut and len(cut_b) == ncut, g
        assert np.array_equal(cut_a, cut_b), g
        worst = max(worst, float(np.abs(a - b).max()))
    print("%s deferred vs inline: max abs difference %.3g over %d measurements" % (core, worst, ncut), flush=True)
    assert worst < 1e-5
