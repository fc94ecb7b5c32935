"""Task-file front end: frequency mesh, cutoff schedule, lattice/model, options.

Reads the reference's XML task schema unchanged (``src/TaskFileParser.cpp:21-347``)
and produces the grids exactly as the reference generates them, in single
precision (the FP64 flow promotes these same float values, it never recomputes
them in double):

* exponential mesh   ``step = powf(max/min, 1/(count-1))``, ``w_i = min*powf(step, i)``
  (``src/TaskFileParser.cpp:107-127``), evaluated with the C library's ``powf``;
* manual mesh        sorted, duplicates rejected (``:128-149``);
* exponential cutoff ``while (max > min) { push(max); max *= step; }`` in float
  (``:157-180``); manual cutoffs sorted descending (``:181-197``);
* model children that are not coupling parameters become core options
  (``:211-220, 313-317``).
"""
from __future__ import annotations

import ctypes
import ctypes.util
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .lattice import (Lattice, LatticeUnitCell, SpinModelUnitCell, new_lattice_model,
                      string_to_float)

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.powf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]


class TaskError(ValueError):
    """Raised where the reference throws ``Exception::Type::InitializationError``."""


def _f32(x) -> np.float32:
    return np.float32(x)


def exponential_mesh(fmin: float, fmax: float, count: float) -> np.ndarray:
    """Positive Matsubara mesh, float32 (``src/TaskFileParser.cpp:112-123``)."""
    fmin, fmax, count = _f32(fmin), _f32(fmax), _f32(count)
    if fmin <= 0:
        raise TaskError("Parameter 'task.parameters.frequency.min' must be positive")
    if fmax <= fmin:
        raise TaskError("Parameter 'task.parameters.frequency.max' must be greater than 'min'")
    if count <= 1:
        raise TaskError("Parameter 'task.parameters.frequency.count' must be greater than 2.")
    ratio = _f32(fmax / fmin)
    expo = _f32(_f32(1.0) / _f32(count - _f32(1.0)))
    step = _f32(_libm.powf(float(ratio), float(expo)))
    out = []
    i = 0
    while i < count:
        out.append(_f32(fmin * _f32(_libm.powf(float(step), float(_f32(i))))))
        i += 1
    return np.array(out, dtype=np.float32)


def manual_mesh(values: List[float]) -> np.ndarray:
    v = sorted(_f32(x) for x in values)
    for a, b in zip(v[:-1], v[1:]):
        if a == b:
            raise TaskError("Duplicate value in task.parameters.frequency.value")
    if len(v) < 2:
        raise TaskError("FrequencyDiscretization must contain at least two frequency values")
    return np.array(v, dtype=np.float32)


def exponential_cutoffs(cmax: float, cmin: float, step: float) -> np.ndarray:
    """Descending cutoff list by repeated float multiplication (``:171-176``)."""
    cmax, cmin, step = _f32(cmax), _f32(cmin), _f32(step)
    if cmin <= 0:
        raise TaskError("Parameter 'task.parameters.cutoff.min' must be positive")
    if cmax <= cmin:
        raise TaskError("Parameter 'task.parameters.cutoff.max' must be greater than 'min'")
    if step <= 0 or step >= 1:
        raise TaskError("Parameter 'task.parameters.cutoff.step' must be in the range (0,1)")
    out = []
    while cmax > cmin:
        out.append(cmax)
        cmax = _f32(cmax * step)
    if len(out) < 2:
        raise TaskError("CutoffDiscretization must contain at least two values")
    return np.array(out, dtype=np.float32)


@dataclass
class MeasurementSpec:
    """``FrgCoreFactory::MeasurementSpecification`` (``src/FrgCoreFactory.hpp:17-29``)."""

    identifier: str
    output: Optional[str] = None
    min_cutoff: float = 0.0
    max_cutoff: float = float("inf")
    defer: bool = False


@dataclass
class Task:
    frequencies: np.ndarray            # positive mesh, float32, ascending
    cutoffs: np.ndarray                # float32, descending
    lattice_name: str
    lattice_range: int
    model_name: str
    symmetry: str                      # core identifier: "SU2" / "TRI" / "XYZ"
    model_options: Dict[str, str]
    measurements: List[MeasurementSpec] = field(default_factory=list)

    def build_lattice(self, resource_path: Optional[str] = None):
        """Run the lattice/model factory; returns ``(lattice, spin_model, core_options)``."""
        uc = LatticeUnitCell.from_resources(self.lattice_name, resource_path)
        model = SpinModelUnitCell.from_resources(self.model_name, self.model_options, resource_path)
        lattice = new_lattice_model(uc, model, self.lattice_range)
        core_options = {k: v for k, v in self.model_options.items() if k not in model.interaction_parameters}
        return lattice, model, core_options


def _text(node: ET.Element, path: str) -> str:
    child = node.find(path)
    if child is None or child.text is None:
        raise TaskError("Invalid task file. Missing '%s'" % path)
    return child.text.strip()


def parse_task(source: str, is_text: bool = False) -> Task:
    root = ET.fromstring(source) if is_text else ET.parse(source).getroot()
    if root.tag != "task":
        raise TaskError("Invalid task file. Root element must be 'task'.")
    params = root.find("parameters")
    if params is None:
        raise TaskError("Invalid task file. Missing 'task.parameters'.")

    fr = params.find("frequency")
    if fr is None:
        raise TaskError("Invalid task file. Missing 'task.parameters.frequency'.")
    kind = fr.get("discretization")
    if kind == "exponential":
        mesh = exponential_mesh(string_to_float(_text(fr, "min")), string_to_float(_text(fr, "max")),
                                string_to_float(_text(fr, "count")))
    elif kind == "manual":
        mesh = manual_mesh([string_to_float(v.text) for v in fr.findall("value")])
    else:
        raise TaskError("Invalid task file. Unknown frequency discretization '%s'" % kind)

    cu = params.find("cutoff")
    if cu is None:
        raise TaskError("Invalid task file. Missing 'task.parameters.cutoff'.")
    kind = cu.get("discretization")
    if kind == "exponential":
        cutoffs = exponential_cutoffs(string_to_float(_text(cu, "max")), string_to_float(_text(cu, "min")),
                                      string_to_float(_text(cu, "step")))
    elif kind == "manual":
        cutoffs = np.array(sorted((string_to_float(v.text) for v in cu.findall("value")), reverse=True),
                           dtype=np.float32)
    else:
        raise TaskError("Invalid task file. Unknown cutoff discretization '%s'" % kind)

    lat = params.find("lattice")
    model = params.find("model")
    if lat is None or model is None or model.get("name") is None or model.get("symmetry") is None:
        raise TaskError("Invalid task file. Lattice or model incomplete.")
    options: Dict[str, str] = {}
    for child in model:
        if child.text is None:
            raise TaskError("Invalid task file. Unspecified parameter value (task.parameters.model.%s)" % child.tag)
        if child.tag in options:
            raise TaskError("Invalid task file. Multiple definitions of parameter '%s'" % child.tag)
        options[child.tag] = child.text.strip()

    measurements: List[MeasurementSpec] = []
    ms = root.find("measurements")
    if ms is not None:
        for m in ms.findall("measurement"):
            spec = MeasurementSpec(identifier=m.get("name"), output=m.get("output"),
                                   defer=(m.get("method") == "defer"))
            if m.get("minCutoff") is not None:
                spec.min_cutoff = float(string_to_float(m.get("minCutoff")))
            if m.get("maxCutoff") is not None:
                spec.max_cutoff = float(string_to_float(m.get("maxCutoff")))
            measurements.append(spec)

    return Task(frequencies=mesh, cutoffs=cutoffs, lattice_name=lat.get("name"),
                lattice_range=int(lat.get("range")), model_name=model.get("name"),
                symmetry=model.get("symmetry"), model_options=options, measurements=measurements)


def make_task_xml(*, lattice: str, lattice_range: int, model: str, symmetry: str,
                  couplings: Dict[str, float], freq_min: float, freq_max: float, freq_count: int,
                  cutoff_max: float, cutoff_min: float, cutoff_step: float,
                  extra_options: Optional[Dict[str, str]] = None) -> str:
    """Write a task file in the reference schema (used for synthetic workloads)."""
    opts = "".join("<%s>%s</%s>" % (k, repr(float(v)) if not isinstance(v, str) else v, k)
                   for k, v in couplings.items())
    if extra_options:
        opts += "".join("<%s>%s</%s>" % (k, v, k) for k, v in extra_options.items())
    return ("<task><parameters>"
            "<frequency discretization=\"exponential\"><min>%r</min><max>%r</max><count>%d</count></frequency>"
            "<cutoff discretization=\"exponential\"><max>%r</max><min>%r</min><step>%r</step></cutoff>"
            "<lattice name=\"%s\" range=\"%d\"/>"
            "<model name=\"%s\" symmetry=\"%s\">%s</model>"
            "</parameters><measurements><measurement name=\"correlation\"/></measurements></task>"
            % (freq_min, freq_max, freq_count, cutoff_max, cutoff_min, cutoff_step,
               lattice, lattice_range, model, symmetry, opts))
