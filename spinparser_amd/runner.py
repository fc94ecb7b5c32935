"""Task lifecycle of a GPU flow: the cutoff loop with checkpoints, deferred
measurements and the post-processing stage, as ``SpinParser::runCore`` drives a
core (``src/SpinParser.cpp:126-220``).

Files follow the reference's fileset (``src/SpinParser.cpp:67-70``): beside the
task file ``name.xml`` live ``name.obs`` (measurements), ``name.checkpoint``
(the state written at checkpoint time and at the end) and ``name.data`` (the
state at every measured cutoff when a measurement is deferred).  The task file
records the computation status in a ``<calculation>`` element
(``src/TaskFileParser.cpp:33-72,349-374``): ``running`` / ``postprocessing`` /
``finished`` with start, checkpoint and end times.

Status ``new`` or ``running`` runs the flow (resuming from the checkpoint file
when running); every ``takeMeasurements`` writes the non-deferred
measurements, and, when any measurement is deferred (``method="defer"``, or
``defer_all`` as the reference's ``--deferMeasurements``), appends the state to
the data file (``src/FrgCore.hpp:36-68``).  The status then becomes
``postprocessing``; the next run replays every state of the data file through
the deferred measurements (``src/SpinParser.cpp:188-204``) and finishes.
"""
from __future__ import annotations

import datetime
import os
import time
import xml.etree.ElementTree as ET
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .frgcore import (HipSU2FrgCore, HipTRIFrgCore, HipXYZFrgCore, InitializationError, new_core_from_task)
from .task import parse_task

NEW, RUNNING, POSTPROCESSING, FINISHED = "new", "running", "postprocessing", "finished"


class IOErrorTask(IOError):
    """``Exception::Type::IOError``."""


@dataclass
class Fileset:
    taskFile: str
    obsFile: str
    dataFile: str
    checkpointFile: str

    @staticmethod
    def of(task_path: str) -> "Fileset":
        base = os.path.splitext(task_path)[0]
        return Fileset(task_path, base + ".obs", base + ".data", base + ".checkpoint")


@dataclass
class ComputationStatus:
    identifier: str = NEW
    startTime: Optional[str] = None
    checkpointTime: Optional[str] = None
    endTime: Optional[str] = None


def timestamp() -> str:
    """Boost's simple time string, e.g. 2024-Jan-05 13:02:11 (``src/lib/Timestamp.hpp:70-73``)."""
    return datetime.datetime.now().strftime("%Y-%b-%d %H:%M:%S")


def read_status(task_path: str) -> ComputationStatus:
    """``TaskFileParser`` status parsing (``src/TaskFileParser.cpp:33-72``)."""
    root = ET.parse(task_path).getroot()
    calc = root.find("calculation")
    if calc is None or calc.get("status") is None:
        return ComputationStatus(NEW)
    st = calc.get("status")
    out = ComputationStatus(st, calc.get("startTime"), calc.get("checkpointTime"), calc.get("endTime"))
    if st not in (RUNNING, POSTPROCESSING, FINISHED):
        raise InitializationError("Checkpoint is corrupted [end time is not specified].")
    if out.startTime is None:
        raise InitializationError("Checkpoint is corrupted [start time is not specified].")
    if st in (POSTPROCESSING, FINISHED) and out.checkpointTime is None:
        raise InitializationError("Checkpoint is corrupted [checkpoint time is not specified].")
    if st == FINISHED and out.endTime is None:
        raise InitializationError("Cannot continue calculation of task. Checkpoint seems corrupted "
                                  "(end time is not specified).")
    return out


def write_task_file(task_path: str, status: ComputationStatus) -> None:
    """``TaskFileParser::writeTaskFile`` (``src/TaskFileParser.cpp:349-374``):
    the previous ``<calculation>`` element is replaced by the current status."""
    tree = ET.parse(task_path)
    root = tree.getroot()
    for old in root.findall("calculation"):
        root.remove(old)
    if status.identifier != NEW:
        calc = ET.SubElement(root, "calculation")
        calc.set("startTime", status.startTime or "")
        calc.set("checkpointTime", status.checkpointTime or "")
        if status.identifier == FINISHED:
            calc.set("endTime", status.endTime or "")
        calc.set("status", status.identifier)
    tree.write(task_path, encoding="utf-8", xml_declaration=True)


def _write_new_records(core, path: str, written: dict) -> None:
    """Append the measurement records taken since the last call to the .obs file."""
    from .obs import write_su2_records, write_tri_records, write_xyz_records
    lattice = core.problem.lattice
    for m in core.measurements():
        n0 = written.get(id(m), 0)
        new = m.records[n0:]
        if not new:
            continue
        if isinstance(core, HipXYZFrgCore):
            write_xyz_records(path, lattice, new)
        elif isinstance(core, HipTRIFrgCore):
            write_tri_records(path, lattice, new)
        else:
            write_su2_records(path, lattice, new)
        written[id(m)] = len(m.records)


def run_task(task_path: str, resource_path: Optional[str] = None, defer_all: bool = False,
             force_restart: bool = False, checkpoint_seconds: float = 3600.0) -> str:
    """One invocation of the reference's executable on a task file (the
    ``SpinParser::run`` / ``runCore`` sequence).  Returns the status after the
    run ("postprocessing" or "finished")."""
    files = Fileset.of(task_path)
    status = ComputationStatus(NEW) if force_restart else read_status(task_path)
    if status.identifier == NEW:
        # no previous checkpoint or data file may survive (src/TaskFileParser.cpp:75-99)
        for path in (files.checkpointFile, files.dataFile):
            if os.path.exists(path):
                if not force_restart:
                    raise IOErrorTask("File [%s] already exists. Specify the --forceRestart option to overwrite "
                                      "existing files." % path)
                os.remove(path)
        if force_restart and os.path.exists(files.obsFile):
            os.remove(files.obsFile)
    if status.identifier == FINISHED:
        return FINISHED

    task = parse_task(task_path)
    core = new_core_from_task(task, resource_path)
    written: dict = {}
    try:
        if status.identifier == POSTPROCESSING:
            # replay every stored state through the deferred measurements
            n = 0
            while core.readCheckpoint(files.dataFile, n, missing_ok=True):
                core.takeMeasurements(postprocessing=True, defer_all=defer_all)
                _write_new_records(core, files.obsFile, written)
                n += 1
            status.endTime = timestamp()
            status.identifier = FINISHED
            write_task_file(task_path, status)
            return FINISHED

        cutoffs = np.asarray(task.cutoffs, dtype=np.float64)
        k = 0
        if status.identifier == RUNNING:
            core.readCheckpoint(files.checkpointFile)
            k = int(np.argmin(np.abs(cutoffs - core.cutoff)))
        else:
            status.startTime = timestamp()
        status.checkpointTime = timestamp()
        last_ckpt = time.monotonic()
        while k < len(cutoffs) - 1:
            core.computeStep()
            core.takeMeasurements(defer_all=defer_all, data_file=files.dataFile)
            _write_new_records(core, files.obsFile, written)
            if core.flow().isDiverged():
                break
            k += 1
            core.finalizeStep(float(cutoffs[k]))
            if time.monotonic() - last_ckpt > checkpoint_seconds:
                last_ckpt = time.monotonic()
                status.checkpointTime = timestamp()
                status.identifier = RUNNING
                core.writeCheckpoint(files.checkpointFile, append=False)
                write_task_file(task_path, status)
        core.takeMeasurements(defer_all=defer_all, data_file=files.dataFile)
        _write_new_records(core, files.obsFile, written)
        deferred = defer_all or any(m.isDeferred() for m in core.measurements())
        if deferred:
            status.identifier = POSTPROCESSING
        else:
            status.endTime = timestamp()
            status.identifier = FINISHED
        status.checkpointTime = timestamp()
        core.writeCheckpoint(files.checkpointFile, append=False)
        write_task_file(task_path, status)
        return status.identifier
    finally:
        core.close()


__all__ = ["Fileset", "ComputationStatus", "read_status", "write_task_file", "run_task", "HipSU2FrgCore"]
