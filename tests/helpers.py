"""Shared builders for the test suite."""
import os

import numpy as np

from spinparser_amd.problem import su2_problem
from spinparser_amd.task import make_task_xml, parse_task

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
GOLDEN_RES = os.path.join(GOLDEN, "res")

REF1_FREQS = "0.31812 0.36329 0.41812 0.46329 0.51334 0.56880 0.63024 0.69833 0.77378 0.85737 0.95 1.0 3.0 10.0"


def manual_task_xml(lattice, lattice_range, model, couplings, symmetry="SU2", freqs=REF1_FREQS,
                    cmax="10", cmin="0.3", cstep="0.9"):
    """Task file of the reference's scripted tests (test/scripted/test_reference1.sh:17-50)."""
    vals = "".join("<value>%s</value>" % v for v in freqs.split())
    opts = "".join("<%s>%s</%s>" % (k, v, k) for k, v in couplings.items())
    return ("<task><parameters><frequency discretization=\"manual\">%s</frequency>"
            "<cutoff discretization=\"exponential\"><max>%s</max><min>%s</min><step>%s</step></cutoff>"
            "<lattice name=\"%s\" range=\"%d\"/><model name=\"%s\" symmetry=\"%s\">%s</model>"
            "</parameters><measurements><measurement name=\"correlation\"/></measurements></task>"
            % (vals, cmax, cmin, cstep, lattice, lattice_range, model, symmetry, opts))


def ref1_problem(j="1.0"):
    """SU2 hyperkagome range 3 (the golden flow test_reference1)."""
    return su2_problem(parse_task(manual_task_xml("hyperkagome", 3, "hyperkagome-heisenberg", {"j": j}), True),
                       GOLDEN_RES)


def python_obs_problem():
    """SU2 honeycomb range 3 (test/scripted/test_pythonObs.sh:16-50)."""
    return su2_problem(parse_task(manual_task_xml("honeycomb", 3, "honeycomb-heisenberg", {"j": "1.0"}), True),
                       GOLDEN_RES)


def exp_problem(lattice, lattice_range, model, couplings, nw, fmin, fmax, cmax, cmin, cstep, spin=None, extra=None):
    extra = dict(extra or {})
    if spin is not None:
        extra["spin"] = spin
    xml = make_task_xml(lattice=lattice, lattice_range=lattice_range, model=model, symmetry="SU2",
                        couplings=couplings, freq_min=fmin, freq_max=fmax, freq_count=nw, cutoff_max=cmax,
                        cutoff_min=cmin, cutoff_step=cstep, extra_options=extra or None)
    return su2_problem(parse_task(xml, is_text=True))


def cfg1_problem(nw=32):
    """Square lattice J1=1, range 4, log mesh [0.005, 50], cutoff 50 -> 0.01 x0.98."""
    return exp_problem("square", 4, "square-heisenberg", {"j": 1.0}, nw, 0.005, 50.0, 50.0, 0.01, 0.98)


def cfg2_problem(nw=64, lattice_range=6):
    """Simple cubic J1=1, J2=0.3, range 6, log mesh [0.002, 100], cutoff 100 -> 0.005 x0.985."""
    return exp_problem("cubic", lattice_range, "cubic-j1j2heisenberg", {"j1": 1.0, "j2": 0.3}, nw, 0.002, 100.0,
                       100.0, 0.005, 0.985)


def max_rel(a, b):
    """Normwise relative error max|a-b| / max|b| (0 if both vanish)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    if den == 0.0:
        return num
    return num / den


def elem_rel(a, b, floor=1e-4):
    """Elementwise relative error max_i |a_i - b_i| / max(|b_i|, floor * max|b|):
    entries below the floor are held to the bar in units of the floor."""
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    top = np.abs(b).max()
    if top == 0.0:
        return float(np.abs(a).max())
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor * top)).max())


def tri_ref3_problem():
    """TRI honeycomb range 3, Heisenberg-Kitaev-Gamma j=0.2 k=1 g=-0.1 (test/scripted/test_reference3.sh:17-60)."""
    from spinparser_amd.problem import tri_problem
    xml = manual_task_xml("honeycomb", 3, "honeycomb-kitaev-gamma", {"j": "0.2", "k": "1.0", "g": "-0.1"},
                          symmetry="TRI")
    return tri_problem(parse_task(xml, True), GOLDEN_RES)


def tri_problem_of(lattice, lattice_range, model, couplings, nw, fmin, fmax, cmax, cmin, cstep, extra=None):
    from spinparser_amd.problem import tri_problem
    xml = make_task_xml(lattice=lattice, lattice_range=lattice_range, model=model, symmetry="TRI",
                        couplings=couplings, freq_min=fmin, freq_max=fmax, freq_count=nw, cutoff_max=cmax,
                        cutoff_min=cmin, cutoff_step=cstep, extra_options=extra)
    return tri_problem(parse_task(xml, is_text=True))


def cfg4_problem(nw=48, lattice_range=5):
    """Kagome J=1 with out-of-plane DM D=0.2, range 5, log mesh [0.005, 50], cutoff 50 -> 0.01 x0.98."""
    return tri_problem_of("kagome", lattice_range, "kagome-DM", {"j": 1.0, "d": 0.2}, nw, 0.005, 50.0, 50.0, 0.01,
                          0.98)


def xyz_ref2_problem(symmetry="XYZ"):
    """Hyperhoneycomb range 3, Heisenberg-Kitaev j=0.1 k=-1 (test/scripted/test_reference2.sh:17-53);
    symmetry "TRI" gives the TRI tables of the same task (the reference's cross-core check)."""
    from spinparser_amd.problem import tri_problem, xyz_problem
    xml = manual_task_xml("hyperhoneycomb", 3, "hyperhoneycomb-kitaev", {"j": "0.1", "k": "-1.0"},
                          symmetry=symmetry)
    build = xyz_problem if symmetry == "XYZ" else tri_problem
    return build(parse_task(xml, True), GOLDEN_RES)
