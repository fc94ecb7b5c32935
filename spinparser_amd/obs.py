"""Observable and checkpoint files in the reference's HDF5 layouts (ctypes over
``libspobs.so``).

``write_correlation`` appends one measurement of one observable group, exactly
as ``SU2MeasurementCorrelation::_writeOutfileCorrelation`` /
``XYZMeasurementCorrelation::_writeOutfileCorrelation`` /
``TRIMeasurementCorrelation::_writeOutfileCorrelation`` do
(``src/SU2/SU2MeasurementCorrelation.cpp:179-355``,
``src/XYZ/XYZMeasurementCorrelation.cpp:214-390``,
``src/TRI/TRIMeasurementCorrelation.cpp:318-493``): meta group with the lattice
vectors, basis and site positions, then ``data/measurement_<k>`` with the
float ``cutoff`` attribute and the (basis, sites in range) float data; a
measurement at a cutoff already in the file is discarded.

``write_checkpoint`` / ``read_checkpoint`` follow ``writeCheckpoint`` /
``readCheckpoint`` of the effective actions (``src/SU2/SU2EffectiveAction.hpp:79-205``,
``src/TRI/TRIEffectiveAction.hpp:79-205``): ``checkpoint_<k>`` groups holding
the cutoff and the float vertex arrays.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional, Tuple

import numpy as np

from .build import OBS_PATH

_f32p = ctypes.POINTER(ctypes.c_float)

SP_OBS_OK, SP_OBS_DUPLICATE = 0, 1


class ObsError(IOError):
    """The writer returned an error (the reference throws ``Exception::Type::IOError``)."""


_lib = None


def load_obs_library() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(OBS_PATH):
            raise ObsError("observable writer %s is not built (needs the HDF5 C library)" % OBS_PATH)
        lib = ctypes.CDLL(OBS_PATH)
        lib.sp_obs_write_correlation.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_float, ctypes.c_int,
                                                 _f32p, ctypes.c_int, _f32p, ctypes.c_int, _f32p, _f32p]
        lib.sp_obs_count.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        lib.sp_obs_read_correlation.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _f32p, _f32p,
                                                ctypes.c_int]
        lib.sp_obs_last_error.restype = ctypes.c_char_p
        _cpp = ctypes.POINTER(ctypes.c_char_p)
        _llp = ctypes.POINTER(ctypes.c_longlong)
        lib.sp_ckpt_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, _cpp, _llp,
                                      ctypes.POINTER(_f32p)]
        lib.sp_ckpt_read.argtypes = [ctypes.c_char_p, ctypes.c_int, _f32p, ctypes.c_int, _cpp, _llp,
                                     ctypes.POINTER(_f32p)]
        lib.sp_ckpt_count.argtypes = [ctypes.c_char_p]
        _lib = lib
    return _lib


def lattice_geometry(lattice) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(lattice vectors, basis positions, positions of the sites in range of
    each basis site) as float32, the meta datasets of an observable group."""
    bravais = np.array([list(v) for v in lattice.bravais], dtype=np.float32)
    basis = np.array([lattice.position(i) for i in lattice.basis_ids], dtype=np.float32)
    sites = np.array([[lattice.position(j) for j in rng] for rng in lattice.ranges], dtype=np.float32)
    return bravais, basis, sites


def _p(a):
    return a.ctypes.data_as(_f32p)


def write_correlation(path: str, group: str, cutoff: float, lattice, data: np.ndarray) -> bool:
    """Append one measurement; returns False if the cutoff was already present."""
    lib = load_obs_library()
    bravais, basis, sites = lattice_geometry(lattice)
    data = np.ascontiguousarray(data, dtype=np.float32)
    B, Rb = sites.shape[0], sites.shape[1]
    if data.shape != (B, Rb):
        raise ObsError("correlation data has shape %s, expected %s" % (data.shape, (B, Rb)))
    rc = lib.sp_obs_write_correlation(path.encode(), group.encode(), ctypes.c_float(float(np.float32(cutoff))),
                                      len(bravais), _p(bravais), B, _p(basis), Rb, _p(np.ascontiguousarray(sites)),
                                      _p(data))
    if rc < 0:
        raise ObsError(lib.sp_obs_last_error().decode())
    return rc == SP_OBS_OK


def read_correlations(path: str, group: str, shape: Tuple[int, int]) -> Tuple[np.ndarray, np.ndarray]:
    """All measurements of a group: (cutoffs[n], data[n, B, Rb])."""
    lib = load_obs_library()
    n = lib.sp_obs_count(path.encode(), group.encode())
    if n < 0:
        raise ObsError(lib.sp_obs_last_error().decode())
    cut = np.zeros(n, dtype=np.float32)
    out = np.zeros((n,) + tuple(shape), dtype=np.float32)
    for k in range(n):
        c = ctypes.c_float(0)
        buf = np.zeros(shape, dtype=np.float32)
        rc = lib.sp_obs_read_correlation(path.encode(), group.encode(), k, ctypes.byref(c), _p(buf), buf.size)
        if rc < 0:
            raise ObsError(lib.sp_obs_last_error().decode())
        cut[k] = c.value
        out[k] = buf
    return cut, out


def write_su2_records(path: str, lattice, records: Iterable) -> int:
    """SU2CorZZ / SU2CorDD groups from ``CorrelationRecord`` objects; returns the
    number of measurements written."""
    n = 0
    for r in records:
        a = write_correlation(path, "SU2CorZZ", r.cutoff, lattice, r.zz)
        b = write_correlation(path, "SU2CorDD", r.cutoff, lattice, r.dd)
        n += int(a and b)
    return n


def write_tri_records(path: str, lattice, records: Iterable, groups: Optional[Iterable[str]] = None) -> int:
    """The ten TRICor* groups from ``TRICorrelationRecord`` objects."""
    n = 0
    for r in records:
        ok = True
        for g in (groups or r.data.keys()):
            ok = write_correlation(path, g, r.cutoff, lattice, r.data[g]) and ok
        n += int(ok)
    return n


def write_xyz_records(path: str, lattice, records: Iterable) -> int:
    """The four XYZCor* groups (``src/XYZ/XYZMeasurementCorrelation.cpp:85-96``)
    from records keyed by group name."""
    return write_tri_records(path, lattice, records, ("XYZCorXX", "XYZCorYY", "XYZCorZZ", "XYZCorDD"))


SP_OBS_DUPLICATE_CKPT = -3


def _arrays_args(names, arrays):
    n = len(names)
    c_names = (ctypes.c_char_p * n)(*[nm.encode() for nm in names])
    sizes = (ctypes.c_longlong * n)(*[a.size for a in arrays])
    ptrs = (_f32p * n)(*[_p(a) for a in arrays])
    return c_names, sizes, ptrs


def write_checkpoint(path: str, cutoff: float, arrays: dict, append: bool = True) -> int:
    """Write one checkpoint (values stored as float); returns its id, or -1 if
    a checkpoint at this cutoff exists (``EffectiveAction::writeCheckpoint``)."""
    lib = load_obs_library()
    names = list(arrays)
    data = [np.ascontiguousarray(arrays[k], dtype=np.float32).ravel() for k in names]
    c_names, sizes, ptrs = _arrays_args(names, data)
    rc = lib.sp_ckpt_write(path.encode(), 1 if append else 0, ctypes.c_float(float(np.float32(cutoff))), len(names),
                           c_names, sizes, ptrs)
    if rc == SP_OBS_DUPLICATE_CKPT:
        return -1
    if rc < 0:
        raise ObsError(lib.sp_obs_last_error().decode())
    return rc


def read_checkpoint(path: str, sizes: dict, checkpoint_id: int = -1):
    """Read checkpoint ``checkpoint_id`` (-1: the last); returns (cutoff, {name:
    float32 array}) (``EffectiveAction::readCheckpoint``)."""
    lib = load_obs_library()
    names = list(sizes)
    data = [np.zeros(int(sizes[k]), dtype=np.float32) for k in names]
    c_names, c_sizes, ptrs = _arrays_args(names, data)
    cut = ctypes.c_float(0)
    rc = lib.sp_ckpt_read(path.encode(), int(checkpoint_id), ctypes.byref(cut), len(names), c_names, c_sizes, ptrs)
    if rc < 0:
        raise ObsError(lib.sp_obs_last_error().decode())
    return float(cut.value), dict(zip(names, data))


def checkpoint_count(path: str) -> int:
    """Number of checkpoints in a data file (0 if it does not exist)."""
    rc = load_obs_library().sp_ckpt_count(path.encode())
    if rc < 0:
        raise ObsError(load_obs_library().sp_obs_last_error().decode())
    return rc
