"""ctypes binding of the C ABI declared in ``include/spfrg.h`` (libspfrg.so).

There is no fallback: if the HIP library is missing or a call fails, an
exception is raised.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

from .build import LIB_PATH

_i32p = ctypes.POINTER(ctypes.c_int)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)

SP_OK, SP_EARG, SP_ENOMEM, SP_EHIP, SP_ESTATE = 0, 1, 2, 3, 4
SP_PRECISION_FP64 = 0
SP_PRECISION_MIXED = 1

KERNEL_DGAMMA, KERNEL_DSIGMA, KERNEL_WEIGHTS, KERNEL_EULER = 0, 1, 2, 3

EXPORTED_SYMBOLS = (
    "sp_ctx_create", "sp_ctx_destroy", "sp_set_state", "sp_get_state", "sp_get_flow", "sp_compute_step",
    "sp_finalize_step", "sp_is_diverged", "sp_measure_correlation", "sp_run_flow", "sp_synchronize", "sp_last_error",
    "sp_profile_enable", "sp_profile_get", "sp_ctx_footprint", "sp_group_create", "sp_group_destroy",
    "sp_group_step", "sp_group_measure", "sp_group_last_error", "sp_set_precision", "sp_compute_step_noalloc",
)


class SpSu2Problem(ctypes.Structure):
    _fields_ = [("nw", ctypes.c_int), ("mesh", _f32p), ("L", ctypes.c_int),
                ("site_rid", _i32p), ("inv_rid", _i32p),
                ("ovl_off", _i32p), ("ovl_rid1", _i32p), ("ovl_rid2", _i32p),
                ("R", ctypes.c_int), ("range_fwd", _i32p), ("range_inv", _i32p),
                ("B", ctypes.c_int), ("Rb", ctypes.c_int), ("corr_rid", _i32p),
                ("spin", ctypes.c_double), ("precision", ctypes.c_int)]


class FlowError(RuntimeError):
    """A C-ABI call returned an error status (the reference throws ``Exception``)."""

    def __init__(self, code: int, message: str):
        super().__init__("%s (status %d)" % (message, code))
        self.code = code


_lib = None


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FlowError(SP_EARG, "HIP library %s is not built; run __graft_entry__.build()" % p)
    lib = ctypes.CDLL(p)
    c_ctx = ctypes.c_void_p
    lib.sp_ctx_create.argtypes = [ctypes.POINTER(SpSu2Problem), ctypes.POINTER(c_ctx)]
    lib.sp_ctx_destroy.argtypes = [c_ctx]
    lib.sp_set_state.argtypes = [c_ctx, ctypes.c_double, _f64p, _f64p, _f64p]
    lib.sp_get_state.argtypes = [c_ctx, _f64p, _f64p, _f64p, _f64p]
    lib.sp_set_precision.argtypes = [c_ctx, ctypes.c_int]
    lib.sp_get_flow.argtypes = [c_ctx, _f64p, _f64p, _f64p]
    lib.sp_compute_step.argtypes = [c_ctx]
    lib.sp_finalize_step.argtypes = [c_ctx, ctypes.c_double]
    lib.sp_is_diverged.argtypes = [c_ctx, _i32p]
    lib.sp_measure_correlation.argtypes = [c_ctx, _f64p, _f64p]
    lib.sp_run_flow.argtypes = [c_ctx, _f64p, ctypes.c_int, _f64p, _f64p, _i32p, _i32p]
    lib.sp_synchronize.argtypes = [c_ctx]
    lib.sp_last_error.argtypes = [c_ctx]
    lib.sp_last_error.restype = ctypes.c_char_p
    lib.sp_profile_enable.argtypes = [c_ctx, ctypes.c_int]
    lib.sp_profile_get.argtypes = [c_ctx, ctypes.c_int, _f64p, ctypes.POINTER(ctypes.c_long)]
    lib.sp_ctx_footprint.argtypes = [ctypes.POINTER(SpSu2Problem), ctypes.POINTER(ctypes.c_longlong)]
    c_grp = ctypes.c_void_p
    _pp = ctypes.POINTER(ctypes.c_void_p)
    lib.sp_group_create.argtypes = [ctypes.c_int, ctypes.POINTER(c_grp)]
    lib.sp_group_destroy.argtypes = [c_grp]
    lib.sp_group_step.argtypes = [c_grp, _pp, ctypes.c_int, _f64p, ctypes.POINTER(_f64p), ctypes.POINTER(_f64p), _i32p]
    lib.sp_group_measure.argtypes = [c_grp, _pp, ctypes.c_int, ctypes.POINTER(_f64p), ctypes.POINTER(_f64p)]
    lib.sp_group_last_error.argtypes = [c_grp]
    lib.sp_group_last_error.restype = ctypes.c_char_p
    _declare_tri(lib)
    _declare_xyz(lib)
    if path is None:
        _lib = lib
    return lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class ProblemBuffers:
    """Keeps the numpy arrays alive behind an ``sp_su2_problem`` struct."""

    def __init__(self, prob, precision: int = SP_PRECISION_FP64):
        keep = [np.ascontiguousarray(prob.mesh, dtype=np.float32)]
        for a in (prob.site_rid, prob.inv_rid, prob.ovl_off, prob.ovl_rid1, prob.ovl_rid2, prob.range_fwd,
                  prob.range_inv, prob.corr_rid):
            keep.append(np.ascontiguousarray(a, dtype=np.int32))
        self.keep = keep
        self.struct = SpSu2Problem(prob.nw, _p(keep[0], _f32p), prob.L, _p(keep[1], _i32p), _p(keep[2], _i32p),
                                   _p(keep[3], _i32p), _p(keep[4], _i32p), _p(keep[5], _i32p), len(keep[6]),
                                   _p(keep[6], _i32p), _p(keep[7], _i32p), prob.B, prob.Rb, _p(keep[8], _i32p),
                                   float(prob.spin), precision)


def footprint_bytes(prob, precision: int = SP_PRECISION_FP64) -> int:
    lib = load_library()
    buf = ProblemBuffers(prob, precision)
    out = ctypes.c_longlong(0)
    rc = lib.sp_ctx_footprint(ctypes.byref(buf.struct), ctypes.byref(out))
    if rc != SP_OK:
        raise FlowError(rc, "sp_ctx_footprint failed")
    return int(out.value)


class FlowContext:
    """One resident flow on the GPU (owner of an ``sp_ctx``)."""

    def __init__(self, prob, precision: int = SP_PRECISION_FP64, lib: Optional[ctypes.CDLL] = None):
        self.lib = lib or load_library()
        self.prob = prob
        self.precision = int(precision)
        self._buf = ProblemBuffers(prob, precision)
        self.ctx = ctypes.c_void_p()
        rc = self.lib.sp_ctx_create(ctypes.byref(self._buf.struct), ctypes.byref(self.ctx))
        if rc != SP_OK:
            msg = self.lib.sp_last_error(self.ctx).decode() if self.ctx else "sp_ctx_create failed"
            if self.ctx:
                self.lib.sp_ctx_destroy(self.ctx)
                self.ctx = ctypes.c_void_p()
            raise FlowError(rc, msg)

    def _check(self, rc: int, what: str):
        if rc != SP_OK:
            raise FlowError(rc, "%s: %s" % (what, self.lib.sp_last_error(self.ctx).decode()))

    def close(self):
        if self.ctx:
            self.lib.sp_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, cutoff: float, v2, v4s, v4d):
        v2, v4s, v4d = (np.ascontiguousarray(x, dtype=np.float64) for x in (v2, v4s, v4d))
        n = self.prob.vertex_size
        if v2.shape != (self.prob.nw,) or v4s.shape != (n,) or v4d.shape != (n,):
            raise FlowError(SP_EARG, "state arrays have the wrong size")
        self._check(self.lib.sp_set_state(self.ctx, float(cutoff), _p(v2, _f64p), _p(v4s, _f64p), _p(v4d, _f64p)),
                    "sp_set_state")

    def set_precision(self, precision: int):
        """Evaluate later derivatives in the FP64 or the mixed form (the state
        stays FP64); the mixed form needs a context created mixed."""
        self._check(self.lib.sp_set_precision(self.ctx, int(precision)), "sp_set_precision")
        self.precision = int(precision)

    def get_state(self) -> Tuple[float, np.ndarray, np.ndarray, np.ndarray]:
        c = ctypes.c_double(0)
        v2 = np.zeros(self.prob.nw)
        v4s = np.zeros(self.prob.vertex_size)
        v4d = np.zeros(self.prob.vertex_size)
        self._check(self.lib.sp_get_state(self.ctx, ctypes.byref(c), _p(v2, _f64p), _p(v4s, _f64p),
                                          _p(v4d, _f64p)), "sp_get_state")
        return c.value, v2, v4s, v4d

    def get_flow(self):
        dv2 = np.zeros(self.prob.nw)
        dss = np.zeros(self.prob.vertex_size)
        ddd = np.zeros(self.prob.vertex_size)
        self._check(self.lib.sp_get_flow(self.ctx, _p(dv2, _f64p), _p(dss, _f64p), _p(ddd, _f64p)), "sp_get_flow")
        return dv2, dss, ddd

    def compute_step(self):
        self._check(self.lib.sp_compute_step(self.ctx), "sp_compute_step")

    def finalize_step(self, new_cutoff: float):
        self._check(self.lib.sp_finalize_step(self.ctx, float(new_cutoff)), "sp_finalize_step")

    def is_diverged(self) -> bool:
        d = ctypes.c_int(0)
        self._check(self.lib.sp_is_diverged(self.ctx, ctypes.byref(d)), "sp_is_diverged")
        return bool(d.value)

    def measure_correlation(self):
        n = self.prob.B * self.prob.Rb
        zz = np.zeros(n)
        dd = np.zeros(n)
        self._check(self.lib.sp_measure_correlation(self.ctx, _p(zz, _f64p), _p(dd, _f64p)), "sp_measure_correlation")
        return zz.reshape(self.prob.B, self.prob.Rb), dd.reshape(self.prob.B, self.prob.Rb)

    def run_flow(self, cutoffs):
        """The cutoff loop over ``cutoffs`` (the state at cutoffs[0]) queued by
        ``sp_run_flow``: returns the records [(zz, dd)] of the computed steps and
        whether the loop stopped on a diverged flow."""
        cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
        n = self.prob.B * self.prob.Rb
        steps = max(len(cut) - 1, 0)
        zz = np.zeros(max(steps, 1) * n)
        dd = np.zeros(max(steps, 1) * n)
        nrec = ctypes.c_int(0)
        div = ctypes.c_int(0)
        self._check(self.lib.sp_run_flow(self.ctx, _p(cut, _f64p), len(cut), _p(zz, _f64p), _p(dd, _f64p),
                                         ctypes.byref(nrec), ctypes.byref(div)), "sp_run_flow")
        shape = (self.prob.B, self.prob.Rb)
        recs = [(zz[k * n:(k + 1) * n].reshape(shape), dd[k * n:(k + 1) * n].reshape(shape)) for k in range(nrec.value)]
        return recs, bool(div.value)

    def synchronize(self):
        self._check(self.lib.sp_synchronize(self.ctx), "sp_synchronize")

    def profile(self, enable: bool):
        self._check(self.lib.sp_profile_enable(self.ctx, 1 if enable else 0), "sp_profile_enable")

    def profile_get(self, kernel: int) -> Tuple[float, int]:
        ms = ctypes.c_double(0)
        n = ctypes.c_long(0)
        self._check(self.lib.sp_profile_get(self.ctx, kernel, ctypes.byref(ms), ctypes.byref(n)), "sp_profile_get")
        return ms.value, n.value


class FlowGroup:
    """Batch of resident flows advanced by one launch per kernel and step
    (``sp_group_*``)."""

    def __init__(self, capacity: int):
        self.lib = load_library()
        self.grp = ctypes.c_void_p()
        rc = self.lib.sp_group_create(int(capacity), ctypes.byref(self.grp))
        if rc != SP_OK:
            raise FlowError(rc, "sp_group_create failed")
        self.capacity = capacity

    def close(self):
        if self.grp:
            self.lib.sp_group_destroy(self.grp)
            self.grp = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != SP_OK:
            raise FlowError(rc, "%s: %s" % (what, self.lib.sp_group_last_error(self.grp).decode()))

    def step(self, flows, new_cutoffs):
        """One cutoff-loop iteration for every flow; returns ([(zz, dd)], diverged[])."""
        n = len(flows)
        ctxs = (ctypes.c_void_p * n)(*[f.ctx.value for f in flows])
        nc = np.ascontiguousarray(new_cutoffs, dtype=np.float64)
        outs = [(np.zeros(f.prob.B * f.prob.Rb), np.zeros(f.prob.B * f.prob.Rb)) for f in flows]
        zz = (_f64p * n)(*[_p(o[0], _f64p) for o in outs])
        dd = (_f64p * n)(*[_p(o[1], _f64p) for o in outs])
        div = np.zeros(n, dtype=np.int32)
        self._check(self.lib.sp_group_step(self.grp, ctxs, n, _p(nc, _f64p), zz, dd, _p(div, _i32p)), "sp_group_step")
        return [(o[0].reshape(f.prob.B, f.prob.Rb), o[1].reshape(f.prob.B, f.prob.Rb)) for o, f in zip(outs, flows)], \
            div.astype(bool)

    def measure(self, flows):
        n = len(flows)
        ctxs = (ctypes.c_void_p * n)(*[f.ctx.value for f in flows])
        outs = [(np.zeros(f.prob.B * f.prob.Rb), np.zeros(f.prob.B * f.prob.Rb)) for f in flows]
        zz = (_f64p * n)(*[_p(o[0], _f64p) for o in outs])
        dd = (_f64p * n)(*[_p(o[1], _f64p) for o in outs])
        self._check(self.lib.sp_group_measure(self.grp, ctxs, n, zz, dd), "sp_group_measure")
        return [(o[0].reshape(f.prob.B, f.prob.Rb), o[1].reshape(f.prob.B, f.prob.Rb)) for o, f in zip(outs, flows)]


# ---------------------------------------------------------------------------
# TRI core (sp_tri_*)
# ---------------------------------------------------------------------------
TRI_EXPORTED_SYMBOLS = (
    "sp_tri_ctx_create", "sp_tri_ctx_destroy", "sp_tri_set_state", "sp_tri_get_state", "sp_tri_get_flow",
    "sp_tri_compute_step", "sp_tri_finalize_step", "sp_tri_is_diverged", "sp_tri_measure_correlation",
    "sp_tri_synchronize", "sp_tri_last_error", "sp_tri_profile_enable", "sp_tri_profile_get",
    "sp_tri_ctx_footprint",
)
TRI_GROUPS = ("TRICorDD", "TRICorXX", "TRICorXY", "TRICorXZ", "TRICorYX", "TRICorYY", "TRICorYZ", "TRICorZX",
              "TRICorZY", "TRICorZZ")


def _declare_tri(lib):
    c_ctx = ctypes.c_void_p
    lib.sp_tri_ctx_create.argtypes = [ctypes.POINTER(SpTriProblem), ctypes.POINTER(c_ctx)]
    lib.sp_tri_ctx_destroy.argtypes = [c_ctx]
    lib.sp_tri_set_state.argtypes = [c_ctx, ctypes.c_double, _f64p, _f64p]
    lib.sp_tri_get_state.argtypes = [c_ctx, _f64p, _f64p, _f64p]
    lib.sp_tri_get_flow.argtypes = [c_ctx, _f64p, _f64p]
    lib.sp_tri_compute_step.argtypes = [c_ctx]
    lib.sp_tri_finalize_step.argtypes = [c_ctx, ctypes.c_double]
    lib.sp_tri_is_diverged.argtypes = [c_ctx, _i32p]
    lib.sp_tri_measure_correlation.argtypes = [c_ctx, _f64p]
    lib.sp_tri_synchronize.argtypes = [c_ctx]
    lib.sp_tri_last_error.argtypes = [c_ctx]
    lib.sp_tri_last_error.restype = ctypes.c_char_p
    lib.sp_tri_profile_enable.argtypes = [c_ctx, ctypes.c_int]
    lib.sp_tri_profile_get.argtypes = [c_ctx, ctypes.c_int, _f64p, ctypes.POINTER(ctypes.c_long)]
    lib.sp_tri_ctx_footprint.argtypes = [ctypes.POINTER(SpTriProblem), ctypes.POINTER(ctypes.c_longlong)]


class SpTriProblem(ctypes.Structure):
    _fields_ = [("nw", ctypes.c_int), ("mesh", _f32p), ("L", ctypes.c_int),
                ("site_rid", _i32p), ("site_perm", _i32p), ("inv_rid", _i32p), ("inv_perm", _i32p),
                ("ovl_off", _i32p), ("ovl_rid1", _i32p), ("ovl_rid2", _i32p), ("ovl_perm1", _i32p),
                ("ovl_perm2", _i32p), ("R", ctypes.c_int), ("range_fwd", _i32p), ("range_inv", _i32p),
                ("B", ctypes.c_int), ("Rb", ctypes.c_int), ("corr_rid", _i32p), ("corr_perm", _i32p)]


class TriProblemBuffers:
    """Keeps the numpy arrays alive behind an ``sp_tri_problem`` struct."""

    def __init__(self, prob):
        keep = [np.ascontiguousarray(prob.mesh, dtype=np.float32)]
        for a in (prob.site_rid, prob.site_perm, prob.inv_rid, prob.inv_perm, prob.ovl_off, prob.ovl_rid1,
                  prob.ovl_rid2, prob.ovl_perm1, prob.ovl_perm2, prob.range_fwd, prob.range_inv, prob.corr_rid,
                  prob.corr_perm):
            keep.append(np.ascontiguousarray(a, dtype=np.int32))
        self.keep = keep
        k = keep
        self.struct = SpTriProblem(prob.nw, _p(k[0], _f32p), prob.L, _p(k[1], _i32p), _p(k[2], _i32p),
                                   _p(k[3], _i32p), _p(k[4], _i32p), _p(k[5], _i32p), _p(k[6], _i32p),
                                   _p(k[7], _i32p), _p(k[8], _i32p), _p(k[9], _i32p), len(k[10]), _p(k[10], _i32p),
                                   _p(k[11], _i32p), prob.B, prob.Rb, _p(k[12], _i32p), _p(k[13], _i32p))


def tri_footprint_bytes(prob) -> int:
    lib = load_library()
    buf = TriProblemBuffers(prob)
    out = ctypes.c_longlong(0)
    rc = lib.sp_tri_ctx_footprint(ctypes.byref(buf.struct), ctypes.byref(out))
    if rc != SP_OK:
        raise FlowError(rc, "sp_tri_ctx_footprint failed")
    return int(out.value)


class TriFlowContext:
    """One resident TRI flow on the GPU (owner of an ``sp_tri_ctx``)."""

    _api = "sp_tri_"
    groups = TRI_GROUPS

    def _f(self, name):
        return getattr(self.lib, self._api + name)

    def __init__(self, prob, lib: Optional[ctypes.CDLL] = None):
        self.lib = lib or load_library()
        self.prob = prob
        self._buf = TriProblemBuffers(prob)
        self.ctx = ctypes.c_void_p()
        rc = self._f("ctx_create")(ctypes.byref(self._buf.struct), ctypes.byref(self.ctx))
        if rc != SP_OK:
            msg = self._f("last_error")(self.ctx).decode() if self.ctx else self._api + "ctx_create failed"
            if self.ctx:
                self._f("ctx_destroy")(self.ctx)
                self.ctx = ctypes.c_void_p()
            raise FlowError(rc, msg)

    def _check(self, rc: int, what: str):
        if rc != SP_OK:
            raise FlowError(rc, "%s: %s" % (what, self._f("last_error")(self.ctx).decode()))

    def close(self):
        if self.ctx:
            self._f("ctx_destroy")(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, cutoff: float, v2, v4):
        v2, v4 = (np.ascontiguousarray(x, dtype=np.float64) for x in (v2, v4))
        if v2.shape != (self.prob.nw,) or v4.shape != (self.prob.vertex_size,):
            raise FlowError(SP_EARG, "state arrays have the wrong size")
        self._check(self._f("set_state")(self.ctx, float(cutoff), _p(v2, _f64p), _p(v4, _f64p)), self._api + "set_state")

    def get_state(self):
        c = ctypes.c_double(0)
        v2 = np.zeros(self.prob.nw)
        v4 = np.zeros(self.prob.vertex_size)
        self._check(self._f("get_state")(self.ctx, ctypes.byref(c), _p(v2, _f64p), _p(v4, _f64p)),
                    self._api + "get_state")
        return c.value, v2, v4

    def get_flow(self):
        dv2 = np.zeros(self.prob.nw)
        dv4 = np.zeros(self.prob.vertex_size)
        self._check(self._f("get_flow")(self.ctx, _p(dv2, _f64p), _p(dv4, _f64p)), self._api + "get_flow")
        return dv2, dv4

    def compute_step(self):
        self._check(self._f("compute_step")(self.ctx), self._api + "compute_step")

    def finalize_step(self, new_cutoff: float):
        self._check(self._f("finalize_step")(self.ctx, float(new_cutoff)), self._api + "finalize_step")

    def is_diverged(self) -> bool:
        d = ctypes.c_int(0)
        self._check(self._f("is_diverged")(self.ctx, ctypes.byref(d)), self._api + "is_diverged")
        return bool(d.value)

    def measure_correlation(self):
        """(groups, B, Rb) array in the order of ``groups`` (TRI_GROUPS: 10)."""
        ng = len(self.groups)
        n = ng * self.prob.B * self.prob.Rb
        out = np.zeros(n)
        self._check(self._f("measure_correlation")(self.ctx, _p(out, _f64p)), self._api + "measure_correlation")
        return out.reshape(ng, self.prob.B, self.prob.Rb)

    def synchronize(self):
        self._check(self._f("synchronize")(self.ctx), self._api + "synchronize")

    def profile(self, enable: bool):
        self._check(self._f("profile_enable")(self.ctx, 1 if enable else 0), self._api + "profile_enable")

    def profile_get(self, kernel: int) -> Tuple[float, int]:
        ms = ctypes.c_double(0)
        n = ctypes.c_long(0)
        self._check(self._f("profile_get")(self.ctx, kernel, ctypes.byref(ms), ctypes.byref(n)),
                    self._api + "profile_get")
        return ms.value, n.value


# ---------------------------------------------------------------------------
# XYZ core (sp_xyz_*): the TRI core's lattice tables and entry points with
# four vertex components stored as the reference's planes X, Y, Z, D
# ---------------------------------------------------------------------------
XYZ_EXPORTED_SYMBOLS = tuple(n.replace("sp_tri_", "sp_xyz_") for n in TRI_EXPORTED_SYMBOLS)
XYZ_GROUPS = ("XYZCorXX", "XYZCorYY", "XYZCorZZ", "XYZCorDD")


def _declare_xyz(lib):
    for name in TRI_EXPORTED_SYMBOLS:
        src, dst = getattr(lib, name), getattr(lib, name.replace("sp_tri_", "sp_xyz_"))
        dst.argtypes = src.argtypes
        dst.restype = src.restype


def xyz_footprint_bytes(prob) -> int:
    lib = load_library()
    buf = TriProblemBuffers(prob)
    out = ctypes.c_longlong(0)
    rc = lib.sp_xyz_ctx_footprint(ctypes.byref(buf.struct), ctypes.byref(out))
    if rc != SP_OK:
        raise FlowError(rc, "sp_xyz_ctx_footprint failed")
    return int(out.value)


class XyzFlowContext(TriFlowContext):
    """One resident XYZ flow on the GPU (owner of an ``sp_xyz_ctx``); state
    arrays are the four planes X, Y, Z, D of ``XYZProblem.vertex_size``."""

    _api = "sp_xyz_"
    groups = XYZ_GROUPS
