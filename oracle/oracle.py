"""ctypes front end of the CPU oracle (``su2_oracle.cpp``) -- test infrastructure.

Only the tests, the smoke check and the CPU-baseline leg of the benchmark load
this module; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_su2.so")

_i32p = ctypes.POINTER(ctypes.c_int)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


class OrProblem(ctypes.Structure):
    _fields_ = [("nw", ctypes.c_int), ("mesh", _f32p), ("L", ctypes.c_int),
                ("site_rid", _i32p), ("inv_rid", _i32p),
                ("ovl_off", _i32p), ("ovl_rid1", _i32p), ("ovl_rid2", _i32p),
                ("R", ctypes.c_int), ("range_fwd", _i32p), ("range_inv", _i32p),
                ("B", ctypes.c_int), ("Rb", ctypes.c_int), ("corr_rid", _i32p),
                ("spin", ctypes.c_double)]


def host_threads() -> int:
    """OpenMP threads for the oracle: OMP_NUM_THREADS when set, else the CPUs
    this process may run on (not every CPU of the machine)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """Holds a problem descriptor in the C layout plus references to its arrays."""

    def __init__(self, prob):
        self.prob = prob
        self._keep = [np.ascontiguousarray(x) for x in (prob.mesh, prob.site_rid, prob.inv_rid, prob.ovl_off,
                                                          prob.ovl_rid1, prob.ovl_rid2, prob.range_fwd,
                                                          prob.range_inv, prob.corr_rid)]
        k = self._keep
        self.c = OrProblem(prob.nw, _ptr(k[0], _f32p), prob.L, _ptr(k[1], _i32p), _ptr(k[2], _i32p),
                           _ptr(k[3], _i32p), _ptr(k[4], _i32p), _ptr(k[5], _i32p), len(prob.range_fwd),
                           _ptr(k[6], _i32p), _ptr(k[7], _i32p), prob.B, prob.Rb, _ptr(k[8], _i32p),
                           float(prob.spin))
        self.lib = lib()

    def compute_step(self, cutoff, v2, v4s, v4d, dtype=np.float64, nthreads=None):
        """One right-hand-side evaluation; returns (dv2, dv4s, dv4d, calls[S,T,U])."""
        nthreads = nthreads or host_threads()
        p = self.prob
        v2, v4s, v4d = (np.ascontiguousarray(x, dtype=dtype) for x in (v2, v4s, v4d))
        dv2 = np.zeros(p.nw, dtype=dtype)
        dss = np.zeros(p.vertex_size, dtype=dtype)
        ddd = np.zeros(p.vertex_size, dtype=dtype)
        calls = np.zeros(3, dtype=np.int64)
        if dtype == np.float64:
            self.lib.or_su2_compute_step_f64(ctypes.byref(self.c), ctypes.c_double(float(cutoff)),
                                             _ptr(v2, _f64p), _ptr(v4s, _f64p), _ptr(v4d, _f64p), _ptr(dv2, _f64p),
                                             _ptr(dss, _f64p), _ptr(ddd, _f64p), ctypes.c_int(nthreads),
                                             _ptr(calls, _lp))
        else:
            self.lib.or_su2_compute_step_f32(ctypes.byref(self.c), ctypes.c_float(float(cutoff)),
                                             _ptr(v2, _f32p), _ptr(v4s, _f32p), _ptr(v4d, _f32p), _ptr(dv2, _f32p),
                                             _ptr(dss, _f32p), _ptr(ddd, _f32p), ctypes.c_int(nthreads),
                                             _ptr(calls, _lp))
        return dv2, dss, ddd, calls

    def items(self, cutoff, v2, dv2, v4s, v4d, it0, it1, dtype=np.float64, nthreads=None):
        """dGamma for items [it0, it1) only (bounded CPU timing samples)."""
        nthreads = nthreads or host_threads()
        p = self.prob
        dss = np.zeros(p.vertex_size, dtype=dtype)
        ddd = np.zeros(p.vertex_size, dtype=dtype)
        calls = np.zeros(3, dtype=np.int64)
        f = self.lib.or_su2_items_f64 if dtype == np.float64 else self.lib.or_su2_items_f32
        t = _f64p if dtype == np.float64 else _f32p
        ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
        v2, dv2, v4s, v4d = (np.ascontiguousarray(x, dtype=dtype) for x in (v2, dv2, v4s, v4d))
        f(ctypes.byref(self.c), ct(float(cutoff)), _ptr(v2, t), _ptr(dv2, t), _ptr(v4s, t), _ptr(v4d, t),
          ctypes.c_int(it0), ctypes.c_int(it1), _ptr(dss, t), _ptr(ddd, t), ctypes.c_int(nthreads), _ptr(calls, _lp))
        return dss, ddd, calls

    def dsigma(self, cutoff, v2, v4s, v4d):
        """dSigma/dLambda only (FP64)."""
        v2, v4s, v4d = (np.ascontiguousarray(x, dtype=np.float64) for x in (v2, v4s, v4d))
        dv2 = np.zeros(self.prob.nw)
        self.lib.or_su2_dsigma_f64(ctypes.byref(self.c), ctypes.c_double(float(cutoff)), _ptr(v2, _f64p),
                                   _ptr(v4s, _f64p), _ptr(v4d, _f64p), _ptr(dv2, _f64p))
        return dv2

    def correlation(self, cutoff, v2, v4s, v4d, dtype=np.float64):
        p = self.prob
        zz = np.zeros(p.B * p.Rb, dtype=dtype)
        dd = np.zeros(p.B * p.Rb, dtype=dtype)
        v2, v4s, v4d = (np.ascontiguousarray(x, dtype=dtype) for x in (v2, v4s, v4d))
        if dtype == np.float64:
            self.lib.or_su2_correlation_f64(ctypes.byref(self.c), ctypes.c_double(float(cutoff)), _ptr(v2, _f64p),
                                            _ptr(v4s, _f64p), _ptr(v4d, _f64p), _ptr(zz, _f64p), _ptr(dd, _f64p))
        else:
            self.lib.or_su2_correlation_f32(ctypes.byref(self.c), ctypes.c_float(float(cutoff)), _ptr(v2, _f32p),
                                            _ptr(v4s, _f32p), _ptr(v4d, _f32p), _ptr(zz, _f32p), _ptr(dd, _f32p))
        return zz.reshape(p.B, p.Rb), dd.reshape(p.B, p.Rb)

    def run_flow(self, dtype=np.float32, nthreads=None, ncut=None):
        """Full flow over the cutoff schedule; returns (zz[n,B,Rb], dd[n,B,Rb], n_records, state)."""
        nthreads = nthreads or host_threads()
        p = self.prob
        cut = np.ascontiguousarray(p.cutoffs if ncut is None else p.cutoffs[:ncut], dtype=np.float32)
        v2, v4s, v4d = p.initial_state(dtype)
        zz = np.zeros((len(cut), p.B, p.Rb), dtype=dtype)
        dd = np.zeros((len(cut), p.B, p.Rb), dtype=dtype)
        t = _f64p if dtype == np.float64 else _f32p
        f = self.lib.or_su2_run_flow_f64 if dtype == np.float64 else self.lib.or_su2_run_flow_f32
        n = f(ctypes.byref(self.c), ctypes.c_int(len(cut)), _ptr(cut, _f32p), _ptr(v2, t), _ptr(v4s, t),
              _ptr(v4d, t), _ptr(zz, t), _ptr(dd, t), ctypes.c_int(nthreads))
        return zz, dd, n, (v2, v4s, v4d)

    @staticmethod
    def integrate_test(mesh, kind, a, b):
        mesh = np.ascontiguousarray(mesh, dtype=np.float32)
        out = np.zeros(2, dtype=np.float32)
        lib().or_integrate_test_f32(ctypes.c_int(len(mesh)), _ptr(mesh, _f32p), ctypes.c_int(kind),
                                    ctypes.c_float(a), ctypes.c_float(b), _ptr(out, _f32p))
        return out


# ---------------------------------------------------------------------------
# TRI core (tri_oracle.cpp)
# ---------------------------------------------------------------------------
TRI_LIB_PATH = os.path.join(HERE, "liboracle_tri.so")
TRI_GROUPS = ("TRICorDD", "TRICorXX", "TRICorXY", "TRICorXZ", "TRICorYX", "TRICorYY", "TRICorYZ", "TRICorZX",
              "TRICorZY", "TRICorZZ")


class OrTriProblem(ctypes.Structure):
    _fields_ = [("nw", ctypes.c_int), ("mesh", _f32p), ("L", ctypes.c_int),
                ("site_rid", _i32p), ("site_perm", _i32p), ("inv_rid", _i32p), ("inv_perm", _i32p),
                ("ovl_off", _i32p), ("ovl_rid1", _i32p), ("ovl_rid2", _i32p), ("ovl_perm1", _i32p),
                ("ovl_perm2", _i32p), ("R", ctypes.c_int), ("range_fwd", _i32p), ("range_inv", _i32p),
                ("B", ctypes.c_int), ("Rb", ctypes.c_int), ("corr_rid", _i32p), ("corr_perm", _i32p)]


_tri_lib = None


def tri_lib():
    global _tri_lib
    if _tri_lib is None:
        if not os.path.exists(TRI_LIB_PATH):
            build()
        _tri_lib = ctypes.CDLL(TRI_LIB_PATH)
    return _tri_lib


class TriOracle:
    """CPU TRI flow (16 spin components) on a ``TRIProblem``."""

    def __init__(self, prob):
        self.prob = prob
        arrs = (prob.mesh.astype(np.float32), prob.site_rid, prob.site_perm, prob.inv_rid, prob.inv_perm,
                prob.ovl_off, prob.ovl_rid1, prob.ovl_rid2, prob.ovl_perm1, prob.ovl_perm2, prob.range_fwd,
                prob.range_inv, prob.corr_rid, prob.corr_perm)
        self._keep = [np.ascontiguousarray(a) if a.dtype == np.float32 else np.ascontiguousarray(a, dtype=np.int32)
                      for a in arrs]
        k = self._keep
        self.c = OrTriProblem(prob.nw, _ptr(k[0], _f32p), prob.L, _ptr(k[1], _i32p), _ptr(k[2], _i32p),
                              _ptr(k[3], _i32p), _ptr(k[4], _i32p), _ptr(k[5], _i32p), _ptr(k[6], _i32p),
                              _ptr(k[7], _i32p), _ptr(k[8], _i32p), _ptr(k[9], _i32p), len(prob.range_fwd),
                              _ptr(k[10], _i32p), _ptr(k[11], _i32p), prob.B, prob.Rb, _ptr(k[12], _i32p),
                              _ptr(k[13], _i32p))
        self.lib = self._load()

    # per-core constants: library, symbol prefix, correlation groups
    _prefix = "or_tri_"
    groups = TRI_GROUPS

    @staticmethod
    def _load():
        return tri_lib()

    def _t(self, dtype):
        return (_f64p, ctypes.c_double, "f64") if dtype == np.float64 else (_f32p, ctypes.c_float, "f32")

    def compute_step(self, cutoff, v2, v4, dtype=np.float64, nthreads=None):
        """(dv2, dv4, calls[S,T,U]) at one cutoff."""
        nthreads = nthreads or host_threads()
        p = self.prob
        t, ct, sfx = self._t(dtype)
        v2, v4 = (np.ascontiguousarray(x, dtype=dtype) for x in (v2, v4))
        dv2 = np.zeros(p.nw, dtype=dtype)
        dv4 = np.zeros(p.vertex_size, dtype=dtype)
        calls = np.zeros(3, dtype=np.int64)
        getattr(self.lib, self._prefix + "compute_step_" + sfx)(ctypes.byref(self.c), ct(float(cutoff)), _ptr(v2, t),
                                                         _ptr(v4, t), _ptr(dv2, t), _ptr(dv4, t),
                                                         ctypes.c_int(nthreads), _ptr(calls, _lp))
        return dv2, dv4, calls

    def items(self, cutoff, v2, dv2, v4, it0, it1, dtype=np.float64, nthreads=None):
        nthreads = nthreads or host_threads()
        t, ct, sfx = self._t(dtype)
        v2, dv2, v4 = (np.ascontiguousarray(x, dtype=dtype) for x in (v2, dv2, v4))
        dv4 = np.zeros(self.prob.vertex_size, dtype=dtype)
        calls = np.zeros(3, dtype=np.int64)
        getattr(self.lib, self._prefix + "items_" + sfx)(ctypes.byref(self.c), ct(float(cutoff)), _ptr(v2, t), _ptr(dv2, t),
                                                  _ptr(v4, t), ctypes.c_int(it0), ctypes.c_int(it1), _ptr(dv4, t),
                                                  ctypes.c_int(nthreads), _ptr(calls, _lp))
        return dv4, calls

    def dsigma(self, cutoff, v2, v4):
        """dSigma/dLambda only (FP64; TRI core)."""
        v2, v4 = (np.ascontiguousarray(x, dtype=np.float64) for x in (v2, v4))
        dv2 = np.zeros(self.prob.nw)
        getattr(self.lib, self._prefix + "dsigma_f64")(ctypes.byref(self.c), ctypes.c_double(float(cutoff)),
                                                       _ptr(v2, _f64p), _ptr(v4, _f64p), _ptr(dv2, _f64p))
        return dv2

    def correlation(self, cutoff, v2, v4, dtype=np.float64):
        """dict group -> (B, Rb) array of the correlation outputs (10 for TRI)."""
        p = self.prob
        t, ct, sfx = self._t(dtype)
        v2, v4 = (np.ascontiguousarray(x, dtype=dtype) for x in (v2, v4))
        ng = len(self.groups)
        out = np.zeros(ng * p.B * p.Rb, dtype=dtype)
        getattr(self.lib, self._prefix + "correlation_" + sfx)(ctypes.byref(self.c), ct(float(cutoff)), _ptr(v2, t),
                                                        _ptr(v4, t), _ptr(out, t))
        out = out.reshape(ng, p.B, p.Rb)
        return {g: out[i] for i, g in enumerate(self.groups)}

    def run_flow(self, dtype=np.float32, nthreads=None, ncut=None):
        """Full flow; returns (records[n, groups, B, Rb], n_records, (v2, v4))."""
        nthreads = nthreads or host_threads()
        p = self.prob
        cut = np.ascontiguousarray(p.cutoffs if ncut is None else p.cutoffs[:ncut], dtype=np.float32)
        v2, v4 = p.initial_state(dtype)
        rec = np.zeros((len(cut), len(self.groups), p.B, p.Rb), dtype=dtype)
        t, ct, sfx = self._t(dtype)
        n = getattr(self.lib, self._prefix + "run_flow_" + sfx)(ctypes.byref(self.c), ctypes.c_int(len(cut)),
                                                         _ptr(cut, _f32p), _ptr(v2, t), _ptr(v4, t), _ptr(rec, t),
                                                         ctypes.c_int(nthreads))
        return rec, n, (v2, v4)

    @staticmethod
    def table(kind):
        """Dense (16, 16, 16) coefficient table: 0 S, 1 U, 2 chalice, 3 inverse chalice, 4 lattice bubble."""
        out = np.zeros(4096)
        n = tri_lib().or_tri_table(ctypes.c_int(kind), _ptr(out, _f64p))
        return out.reshape(16, 16, 16), n


# ---------------------------------------------------------------------------
# XYZ core (xyz_oracle.cpp)
# ---------------------------------------------------------------------------
XYZ_LIB_PATH = os.path.join(HERE, "liboracle_xyz.so")
XYZ_GROUPS = ("XYZCorXX", "XYZCorYY", "XYZCorZZ", "XYZCorDD")

_xyz_lib = None


def xyz_lib():
    global _xyz_lib
    if _xyz_lib is None:
        if not os.path.exists(XYZ_LIB_PATH):
            build()
        _xyz_lib = ctypes.CDLL(XYZ_LIB_PATH)
    return _xyz_lib


class XyzOracle(TriOracle):
    """CPU XYZ flow (4 spin components, state as four planes X, Y, Z, D) on an
    ``XYZProblem``; same lattice tables and entry points as the TRI oracle."""

    _prefix = "or_xyz_"
    groups = XYZ_GROUPS

    @staticmethod
    def _load():
        return xyz_lib()

    @staticmethod
    def table(kind):
        raise NotImplementedError("the XYZ channel algebra has no generated tables")
