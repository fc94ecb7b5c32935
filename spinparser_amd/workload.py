"""Work accounting for one right-hand-side evaluation.

The number of integrand evaluations of a channel with transfer frequency
X = mesh[x] depends only on (x, Lambda): the conventional single-scale terms at
w' = +-Lambda (``src/SU2/SU2FrgCore.cpp:351-371``) plus the trapezoid points of
the three Katanin windows (``:378-424``, ``src/lib/Integrator.hpp:137-287``).
Summing over the items that carry each x gives the exact integrand-call counts
of one evaluation, and with the per-call costs of the survey's work model
(94 L FLOP per s/u kernel call, 106 L + 4 O per t call, an FMA counted as two)
the algorithmic FLOP of the evaluation.
"""
from __future__ import annotations

import numpy as np


def _lesser_pos(m, w):
    if w <= m[0]:
        return 0
    for i in range(1, len(m)):
        if m[i] > w:
            return i - 1
    return len(m) - 1


def _greater_pos(m, w):
    if w <= m[0]:
        return 0
    for i in range(1, len(m)):
        if m[i] > w:
            return i
    return len(m) - 1


def _lesser(m, w):
    return -(_greater_pos(m, -w) + 1) if w < 0 else _lesser_pos(m, w)


def _greater(m, w):
    return -(_lesser_pos(m, -w) + 1) if w < 0 else _greater_pos(m, w)


def window_points(m: np.ndarray, X: float, cut: float) -> int:
    """Integrand evaluations of the three Katanin windows of one channel."""
    n = len(m)
    at = lambda i: m[i] if i >= 0 else -m[-i - 1]  # noqa: E731
    cnt = 0
    if -(X + cut) > at(-n):
        umax = _lesser(m, -(X + cut))
        cnt += (umax - (-n) + 2) if umax != -n else 2
    if X - cut > cut:
        umin, umax = _greater(m, cut - X), _lesser(m, -cut)
        cnt += (umax - umin + 3) if umax >= umin else 2
    if cut < at(n - 1):
        umin = _greater(m, cut)
        cnt += (n - 1 - umin + 2) if umin != n - 1 else 2
    return cnt


def channel_calls(mesh: np.ndarray, cut: float) -> np.ndarray:
    """calls[x] = integrand evaluations of a channel with X = mesh[x] (conventional + windows)."""
    m = np.asarray(mesh, dtype=np.float64)
    cut = float(cut)
    out = np.zeros(len(m), dtype=np.int64)
    for x in range(len(m)):
        X = float(m[x])
        out[x] = (2 if X > 2.0 * cut else 1) + window_points(m, X, cut)
    return out


def evaluation_calls(mesh: np.ndarray, cut: float):
    """(calls_S, calls_T, calls_U) summed over all N_f items of one evaluation."""
    nw = len(mesh)
    c = channel_calls(mesh, cut)
    x = np.arange(nw)
    items_with_s = (x + 1) * nw          # u <= s, any t
    items_with_t = np.full(nw, nw * (nw + 1) // 2)
    items_with_u = (nw - x) * nw         # s >= u, any t
    return int((c * items_with_s).sum()), int((c * items_with_t).sum()), int((c * items_with_u).sum())


def evaluation_flops(prob, cut: float) -> float:
    cs, ct, cu = evaluation_calls(prob.mesh, cut)
    return prob.flops_per_evaluation(cs, ct, cu)


def compulsory_bytes(prob, value_bytes: int = 8) -> int:
    """Read Gamma and Sigma, write dGamma and dSigma once per evaluation."""
    return 2 * (2 * prob.vertex_size + prob.nw) * value_bytes
