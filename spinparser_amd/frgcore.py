"""Python mirror of the reference's FrgCore plugin boundary over the C ABI.

Names and argument meaning follow the reference so that a flow is driven the
same way ``SpinParser::runCore`` drives it:

* ``EffectiveAction`` (``src/EffectiveAction.hpp:19-60``): ``cutoff``,
  ``isDiverged()``;
* ``FrgCore`` (``src/FrgCore.hpp:28-141``): ``computeStep()``,
  ``finalizeStep(newCutoff)``, ``takeMeasurements()``, ``flowingFunctional()``,
  ``flow()``;
* ``FrgCoreFactory.newFrgCore(identifier, ...)`` (``src/FrgCoreFactory.cpp:25-52``)
  selects the GPU core for ``symmetry="SU2"``, ``"XYZ"`` or ``"TRI"`` when the
  model carries the option ``<backend>hip</backend>`` (or when ``backend="hip"``
  is passed);
* ``run_core`` restates the cutoff loop of ``src/SpinParser.cpp:126-186``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .capi import FlowContext, SP_PRECISION_FP64, TriFlowContext, XyzFlowContext
from .problem import SU2Problem, TRIProblem, su2_problem, tri_problem, xyz_problem
from .task import MeasurementSpec, Task


class InitializationError(RuntimeError):
    """``Exception::Type::InitializationError``."""


class ArgumentError(ValueError):
    """``Exception::Type::ArgumentError``."""


class EffectiveAction:
    """State view of a GPU-resident flow (``src/EffectiveAction.hpp:19-60``)."""

    def __init__(self, ctx: FlowContext, which: str):
        self._ctx = ctx
        self._which = which  # "state" or "flow"
        self._cutoff = 0.0

    @property
    def cutoff(self) -> float:
        return self._ctx_cutoff() if self._which == "state" else self._cutoff

    def _ctx_cutoff(self) -> float:
        return self._core_cutoff

    def isDiverged(self) -> bool:
        if self._which != "flow":
            raise ArgumentError("isDiverged is evaluated on the flow (FrgCore::_flow)")
        return self._ctx.is_diverged()

    def arrays(self):
        """Download the vertex arrays in the reference layout: (v2, v4ss, v4dd)
        for the SU(2) core, (v2, v4) for the TRI and XYZ cores (XYZ: the four
        planes X, Y, Z, D concatenated)."""
        if self._which == "state":
            return tuple(self._ctx.get_state()[1:])
        return self._ctx.get_flow()


class Measurement:
    """``src/Measurement.hpp``: cutoff window and defer flag."""

    def __init__(self, spec: MeasurementSpec):
        self.spec = spec

    def minCutoff(self) -> float:
        return self.spec.min_cutoff

    def maxCutoff(self) -> float:
        return self.spec.max_cutoff

    def isDeferred(self) -> bool:
        return self.spec.defer


@dataclass
class CorrelationRecord:
    cutoff: float
    zz: np.ndarray   # SU2CorZZ data, (basis, sites in range)
    dd: np.ndarray   # SU2CorDD data


class SU2MeasurementCorrelation(Measurement):
    """GPU correlation measurement (``src/SU2/SU2MeasurementCorrelation.cpp:66-177``).
    Records are kept in memory in the order the reference writes ``measurement_k``
    groups; duplicates at the same cutoff are discarded like the reference does
    (``:318-322``)."""

    def __init__(self, spec: MeasurementSpec):
        super().__init__(spec)
        self.records: List[CorrelationRecord] = []

    def takeMeasurement(self, core: "HipSU2FrgCore", isMasterTask: bool = True) -> None:
        cutoff = core.cutoff
        if self.records and np.float32(self.records[-1].cutoff) == np.float32(cutoff):
            return
        zz, dd = core.context.measure_correlation()
        if isMasterTask:
            self.records.append(CorrelationRecord(cutoff, zz, dd))


@dataclass
class TRICorrelationRecord:
    cutoff: float
    data: Dict[str, np.ndarray]   # TRICorDD, TRICorXX, ... TRICorZZ: (basis, sites in range)


class TRIMeasurementCorrelation(Measurement):
    """GPU correlation measurement of the TRI core
    (``src/TRI/TRIMeasurementCorrelation.cpp:147-316``): the 10 groups
    ``TRICorDD``, ``TRICorXX`` ... ``TRICorZZ``; duplicate cutoffs are discarded
    like the reference does."""

    def __init__(self, spec: MeasurementSpec):
        super().__init__(spec)
        self.records: List[TRICorrelationRecord] = []

    def takeMeasurement(self, core: "HipTRIFrgCore", isMasterTask: bool = True) -> None:
        cutoff = core.cutoff
        if self.records and np.float32(self.records[-1].cutoff) == np.float32(cutoff):
            return
        out = core.context.measure_correlation()
        if isMasterTask:
            groups = core.context.groups
            self.records.append(TRICorrelationRecord(cutoff, {g: out[i] for i, g in enumerate(groups)}))


class XYZMeasurementCorrelation(TRIMeasurementCorrelation):
    """GPU correlation measurement of the XYZ core
    (``src/XYZ/XYZMeasurementCorrelation.cpp:85-212``): the groups ``XYZCorXX``,
    ``XYZCorYY``, ``XYZCorZZ``, ``XYZCorDD`` (records keyed by group name)."""


class FrgCore:
    """Abstract core (``src/FrgCore.hpp:28-141``)."""

    def __init__(self, measurements: List[Measurement]):
        self._measurements = list(measurements)

    def computeStep(self) -> None:
        raise NotImplementedError

    def finalizeStep(self, newCutoff: float) -> None:
        raise NotImplementedError

    def measurements(self) -> List[Measurement]:
        return self._measurements

    def takeMeasurements(self, postprocessing: bool = False, defer_all: bool = False,
                         data_file: Optional[str] = None) -> None:
        """``FrgCore::takeMeasurements`` (src/FrgCore.hpp:36-68).  In the flow
        (postprocessing False): the measurements that are not deferred, inside
        their cutoff window; then, if any measurement is deferred (or
        ``defer_all``, the reference's --deferMeasurements), the state is
        appended to ``data_file`` for the post-processing stage.  In the
        post-processing stage: the deferred measurements on the state read back
        from the data file."""
        deferred = lambda m: defer_all or m.isDeferred()  # noqa: E731
        for m in self._measurements:
            if m.minCutoff() <= self.cutoff <= m.maxCutoff() and deferred(m) == postprocessing:
                m.takeMeasurement(self)
        if not postprocessing and any(deferred(m) for m in self._measurements):
            if data_file is None:
                raise ArgumentError("deferred measurements need a data file for the post-processing stage")
            self.writeCheckpoint(data_file, append=True)


class HipSU2FrgCore(FrgCore):
    """SU(2) core on the GPU: state, flow and tables resident in HBM."""

    def __init__(self, problem: SU2Problem, measurements: List[Measurement], precision: int = SP_PRECISION_FP64):
        super().__init__(measurements)
        self.problem = problem
        self.spinLength = float(problem.spin)
        self.normalization = float(problem.normalization)
        self.context = FlowContext(problem, precision)
        self.cutoff = float(problem.cutoffs[0])
        v2, v4s, v4d = problem.initial_state(np.float64)
        self.context.set_state(self.cutoff, v2, v4s, v4d)
        self._state = EffectiveAction(self.context, "state")
        self._state._core_cutoff = self.cutoff
        self._flow = EffectiveAction(self.context, "flow")

    def computeStep(self) -> None:
        self.context.compute_step()
        self._flow._cutoff = self.cutoff

    def finalizeStep(self, newCutoff: float) -> None:
        self.context.finalize_step(float(newCutoff))
        self.cutoff = float(newCutoff)
        self._state._core_cutoff = self.cutoff

    def flowingFunctional(self) -> EffectiveAction:
        return self._state

    def flow(self) -> EffectiveAction:
        return self._flow

    def setState(self, cutoff: float, v2, v4s, v4d) -> None:
        """``readCheckpoint`` equivalent: replace the resident state."""
        self.context.set_state(cutoff, v2, v4s, v4d)
        self.cutoff = float(cutoff)
        self._state._core_cutoff = self.cutoff

    def writeCheckpoint(self, dataFilePath: str, append: bool = False) -> int:
        """``SU2EffectiveAction::writeCheckpoint`` (src/SU2/SU2EffectiveAction.hpp:79-150):
        the resident state as float datasets cutoff, v2, v4dd, v4ss; returns the
        checkpoint id, or -1 if the file holds one at this cutoff."""
        from .obs import write_checkpoint
        _, v2, v4s, v4d = self.context.get_state()
        return write_checkpoint(dataFilePath, self.cutoff, {"v2": v2, "v4dd": v4d, "v4ss": v4s}, append)

    def readCheckpoint(self, dataFilePath: str, checkpointId: int = -1, missing_ok: bool = False) -> bool:
        """``SU2EffectiveAction::readCheckpoint`` (:159-205); -1 reads the last checkpoint."""
        from .obs import checkpoint_count, read_checkpoint
        if missing_ok and not 0 <= checkpointId < checkpoint_count(dataFilePath):
            return False  # the reference's readCheckpoint returns false past the last checkpoint
        n = self.problem.vertex_size
        cut, a = read_checkpoint(dataFilePath, {"v2": self.problem.nw, "v4dd": n, "v4ss": n}, checkpointId)
        self.setState(cut, a["v2"], a["v4ss"], a["v4dd"])
        return True

    def close(self) -> None:
        self.context.close()


class HipTRIFrgCore(FrgCore):
    """TRI core on the GPU (``src/TRI/TRIFrgCore.cpp``): 16-component vertex
    resident in HBM, FP64."""

    _context_type = TriFlowContext

    def __init__(self, problem: TRIProblem, measurements: List[Measurement]):
        super().__init__(measurements)
        self.problem = problem
        self.normalization = float(problem.normalization)
        self.context = self._context_type(problem)
        self.cutoff = float(problem.cutoffs[0])
        v2, v4 = problem.initial_state(np.float64)
        self.context.set_state(self.cutoff, v2, v4)
        self._state = EffectiveAction(self.context, "state")
        self._state._core_cutoff = self.cutoff
        self._flow = EffectiveAction(self.context, "flow")

    def computeStep(self) -> None:
        self.context.compute_step()
        self._flow._cutoff = self.cutoff

    def finalizeStep(self, newCutoff: float) -> None:
        self.context.finalize_step(float(newCutoff))
        self.cutoff = float(newCutoff)
        self._state._core_cutoff = self.cutoff

    def flowingFunctional(self) -> EffectiveAction:
        return self._state

    def flow(self) -> EffectiveAction:
        return self._flow

    def setState(self, cutoff: float, v2, v4) -> None:
        self.context.set_state(cutoff, v2, v4)
        self.cutoff = float(cutoff)
        self._state._core_cutoff = self.cutoff

    def writeCheckpoint(self, dataFilePath: str, append: bool = False) -> int:
        """``TRIEffectiveAction::writeCheckpoint`` (src/TRI/TRIEffectiveAction.hpp:79-150)."""
        from .obs import write_checkpoint
        _, v2, v4 = self.context.get_state()
        return write_checkpoint(dataFilePath, self.cutoff, {"v2": v2, "v4": v4}, append)

    def readCheckpoint(self, dataFilePath: str, checkpointId: int = -1, missing_ok: bool = False) -> bool:
        """``TRIEffectiveAction::readCheckpoint`` (:159-205)."""
        from .obs import checkpoint_count, read_checkpoint
        if missing_ok and not 0 <= checkpointId < checkpoint_count(dataFilePath):
            return False  # the reference's readCheckpoint returns false past the last checkpoint
        cut, a = read_checkpoint(dataFilePath, {"v2": self.problem.nw, "v4": self.problem.vertex_size}, checkpointId)
        self.setState(cut, a["v2"], a["v4"])
        return True

    def close(self) -> None:
        self.context.close()


class HipXYZFrgCore(HipTRIFrgCore):
    """XYZ core on the GPU (``src/XYZ/XYZFrgCore.cpp``): four-component vertex
    (planes X, Y, Z, D) resident in HBM, FP64."""

    _context_type = XyzFlowContext
    _planes = ("v4xx", "v4yy", "v4zz", "v4dd")

    def writeCheckpoint(self, dataFilePath: str, append: bool = False) -> int:
        """``XYZEffectiveAction::writeCheckpoint`` (src/XYZ/XYZEffectiveAction.hpp:82-155):
        datasets cutoff, v2, v4dd, v4xx, v4yy, v4zz."""
        from .obs import write_checkpoint
        _, v2, v4 = self.context.get_state()
        planes = v4.reshape(4, -1)
        arrays = {"v2": v2, "v4dd": planes[3], "v4xx": planes[0], "v4yy": planes[1], "v4zz": planes[2]}
        return write_checkpoint(dataFilePath, self.cutoff, arrays, append)

    def readCheckpoint(self, dataFilePath: str, checkpointId: int = -1, missing_ok: bool = False) -> bool:
        """``XYZEffectiveAction::readCheckpoint`` (:164-217)."""
        from .obs import checkpoint_count, read_checkpoint
        if missing_ok and not 0 <= checkpointId < checkpoint_count(dataFilePath):
            return False  # the reference's readCheckpoint returns false past the last checkpoint
        n = self.problem.vertex_size // 4
        sizes = {"v2": self.problem.nw, "v4dd": n, "v4xx": n, "v4yy": n, "v4zz": n}
        cut, a = read_checkpoint(dataFilePath, sizes, checkpointId)
        self.setState(cut, a["v2"], np.concatenate([a[k] for k in self._planes]))
        return True


class FrgCoreFactory:
    @staticmethod
    def newFrgCore(identifier: str, problem, measurements: List[MeasurementSpec],
                   options: Optional[Dict[str, str]] = None, backend: Optional[str] = None) -> FrgCore:
        options = dict(options or {})
        backend = backend or options.pop("backend", None)
        objs: List[Measurement] = []
        for spec in measurements:
            if spec.identifier != "correlation":
                raise InitializationError("Measurement: Unknown measurement type '%s'." % spec.identifier)
            if identifier == "SU2":
                objs.append(SU2MeasurementCorrelation(spec))
            elif identifier == "TRI":
                objs.append(TRIMeasurementCorrelation(spec))
            elif identifier == "XYZ":
                objs.append(XYZMeasurementCorrelation(spec))
            else:
                raise InitializationError("Measurement [correlation]: no GPU measurement for '%s'." % identifier)
        if identifier == "SU2" and backend == "hip":
            return HipSU2FrgCore(problem, objs)
        if identifier == "TRI" and backend == "hip":
            return HipTRIFrgCore(problem, objs)
        if identifier == "XYZ" and backend == "hip":
            return HipXYZFrgCore(problem, objs)
        if identifier in ("SU2", "XYZ", "TRI"):
            raise ArgumentError("Core '%s' with backend '%s' is not provided by this engine" % (identifier, backend))
        raise ArgumentError("Spin model identifier '%s' does not exist." % identifier)


def new_core_from_task(task: Task, resource_path: Optional[str] = None, backend: str = "hip") -> FrgCore:
    if task.symmetry == "TRI":
        prob = tri_problem(task, resource_path)
    elif task.symmetry == "XYZ":
        prob = xyz_problem(task, resource_path)
    else:
        prob = su2_problem(task, resource_path)
    options = {k: v for k, v in task.model_options.items() if k in ("spin", "normalization", "backend")}
    return FrgCoreFactory.newFrgCore(task.symmetry, prob, task.measurements or [MeasurementSpec("correlation")],
                                     options, backend=backend)


def run_core(core: FrgCore, cutoffs: np.ndarray) -> bool:
    """The cutoff loop of ``SpinParser::runCore`` (``src/SpinParser.cpp:141-172``).
    Returns False if the flow diverged."""
    diverged = False
    k = 0
    while k < len(cutoffs) - 1:
        core.computeStep()
        core.takeMeasurements()
        if core.flow().isDiverged():
            diverged = True
            break
        k += 1
        core.finalizeStep(float(cutoffs[k]))
    core.takeMeasurements()
    return not diverged
