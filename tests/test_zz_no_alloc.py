"""sp_compute_step_noalloc (the FP64 computeStep with all scratch in a static
device array) against sp_compute_step on a context holding the same state.

Both run the same kernels on the same inputs, so the flows must agree bit for
bit.  Two calls in a row, the second with a smaller problem, show that the
shared scratch array carries nothing from one call into the next.
"""
import ctypes
import sys

import numpy as np
import pytest

from helpers import cfg1_problem, exp_problem

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _declare(lib):
    f64p = ctypes.POINTER(ctypes.c_double)
    from spinparser_amd.capi import SpSu2Problem
    lib.sp_compute_step_noalloc.argtypes = [ctypes.POINTER(SpSu2Problem), ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_int)]
    lib.sp_compute_step_noalloc.restype = ctypes.c_int
    return f64p


def _advanced_state(prob, k):
    """The FP64 flow's own state after k steps (sp_run_flow)."""
    from spinparser_amd.capi import FlowContext
    ctx = FlowContext(prob)
    ctx.set_state(float(prob.cutoffs[0]), *prob.initial_state(np.float64))
    _, div = ctx.run_flow(prob.cutoffs[:k + 1])
    assert not div
    st = ctx.get_state()
    ctx.close()
    return st


def _reference_flow(prob, state):
    from spinparser_amd.capi import FlowContext
    ctx = FlowContext(prob)
    ctx.set_state(*state)
    ctx.compute_step()
    out = ctx.get_flow()
    ctx.close()
    return out


def _hip_runtime():
    """The HIP runtime the library itself runs on (read from this process's
    mapped libraries after the library is loaded), for the test's own device
    buffers and stream."""
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line and "torch" not in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 is not loaded"
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return hip


H2D, D2H = 1, 2


def _noalloc_flow(lib, prob, state):
    from spinparser_amd.capi import ProblemBuffers
    hip = _hip_runtime()
    cut, v2, v4s, v4d = state
    buf = ProblemBuffers(prob)
    host = {"v2": np.ascontiguousarray(v2), "v4": np.ascontiguousarray(np.stack([v4s, v4d], axis=-1).ravel()),
            "dv2": np.full(prob.nw, np.nan), "dv4": np.full(2 * prob.vertex_size, np.nan)}
    dev = {}
    for k, a in host.items():
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        dev[k] = p
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    div = ctypes.c_int(-1)
    rc = lib.sp_compute_step_noalloc(ctypes.byref(buf.struct), float(cut), dev["v2"], dev["v4"], dev["dv2"],
                                     dev["dv4"], stream, ctypes.byref(div))
    assert hip.hipStreamSynchronize(stream) == 0
    assert rc == 0, "sp_compute_step_noalloc returned %d" % rc
    assert div.value == 0
    for k in ("dv2", "dv4"):
        assert hip.hipMemcpy(host[k].ctypes.data, dev[k], host[k].nbytes, D2H) == 0
    assert hip.hipStreamDestroy(stream) == 0
    for p in dev.values():
        assert hip.hipFree(p) == 0
    dv4 = host["dv4"].reshape(-1, 2)
    return host["dv2"], dv4[:, 0].copy(), dv4[:, 1].copy()


def test_noalloc_equals_compute_step_twice():
    from spinparser_amd.capi import load_library
    lib = load_library()
    _declare(lib)
    cases = [("square r4 nw16", cfg1_problem(nw=16), 200),
             ("cubic r3 nw8 (smaller)", exp_problem("cubic", 3, "cubic-j1j2heisenberg", {"j1": 1.0, "j2": 0.5}, 8,
                                                    0.002, 100.0, 100.0, 0.005, 0.985), 300)]
    for name, prob, k in cases:
        _log("no-alloc case %s: L=%d nw=%d cutoff index %d" % (name, prob.L, prob.nw, k))
        state = _advanced_state(prob, k)
        ref = _reference_flow(prob, state)
        got = _noalloc_flow(lib, prob, state)
        for what, a, b in zip(("dSigma", "dGamma spin", "dGamma density"), got, ref):
            assert np.all(np.isfinite(a)), what
            assert np.array_equal(a, b), "%s differs (max abs %.3g)" % (what, np.abs(a - b).max())


def test_noalloc_rejects_mixed_and_oversized():
    from spinparser_amd.capi import SP_PRECISION_MIXED, ProblemBuffers, load_library
    lib = load_library()
    _declare(lib)
    prob = cfg1_problem(nw=8)
    div = ctypes.c_int(0)
    buf = ProblemBuffers(prob, SP_PRECISION_MIXED)
    dummy = ctypes.c_void_p(16)  # never dereferenced: the call rejects the problem first
    assert lib.sp_compute_step_noalloc(ctypes.byref(buf.struct), 1.0, dummy, dummy, dummy, dummy, None,
                                       ctypes.byref(div)) == 1
    # lattice tables beyond the scratch array: 500,000 overlap entries take
    # 5 MB (two int32 lists and the packed 16-bit indices) > 4 MiB
    big = ProblemBuffers(prob)
    n_ovl = 500000
    off = np.zeros(prob.L + 1, dtype=np.int32)
    off[1:] = n_ovl
    zeros = np.zeros(n_ovl, dtype=np.int32)
    keep = (off, zeros)
    i32p = ctypes.POINTER(ctypes.c_int)
    big.struct.ovl_off = off.ctypes.data_as(i32p)
    big.struct.ovl_rid1 = zeros.ctypes.data_as(i32p)
    big.struct.ovl_rid2 = zeros.ctypes.data_as(i32p)
    assert lib.sp_compute_step_noalloc(ctypes.byref(big.struct), 1.0, dummy, dummy, dummy, dummy, None,
                                       ctypes.byref(div)) == 1
    del keep
