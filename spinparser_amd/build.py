"""Build recipe for the HIP library (gfx950) -- compiled in-tree with hipcc.

``python -m spinparser_amd.build`` (or ``__graft_entry__.build()``) writes
``spinparser_amd/_lib/libspfrg.so`` (the flow kernels and their C ABI) and
``spinparser_amd/_lib/libspobs.so`` (host-only HDF5 writer of observable files,
built when the HDF5 C library is found).  Device code is compiled with
``-ffp-contract=off`` so that frequency arguments, window predicates and mesh
bracket comparisons round exactly like the reference; the hot accumulations use
explicit ``fma()``.  Packed-FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libspfrg.so")
SOURCES = [os.path.join(PKG, "csrc", f) for f in ("su2_flow.hip", "tri_flow.hip", "xyz_flow.hip")]
OBS_PATH = os.path.join(LIB_DIR, "libspobs.so")
OBS_SOURCE = os.path.join(PKG, "csrc", "obs_writer.cpp")
HDF5_PREFIXES = ("/opt/conda", "/usr", "/usr/local")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_lib(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-fconstexpr-steps=33554432", "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
           "-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for libspfrg.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def hdf5_prefix():
    for pre in HDF5_PREFIXES:
        if os.path.exists(os.path.join(pre, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(pre, "lib", "libhdf5.so")):
            return pre
    return None


def build_obs(verbose: bool = False):
    """Host-only HDF5 observable writer; returns its path, or None without HDF5."""
    pre = hdf5_prefix()
    if pre is None:
        return None
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = OBS_PATH + ".tmp"
    lib = os.path.join(pre, "lib")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(pre, "include"), "-o", tmp, OBS_SOURCE, "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed for libspobs.so")
    os.replace(tmp, OBS_PATH)
    return OBS_PATH


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_obs(verbose=True))
