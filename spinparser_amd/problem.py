"""SU(2) flow problem descriptor: the flattened tables the flow kernels consume.

Everything the right-hand side needs besides the vertex state, derived once from
the task file and the lattice factory:

* ``mesh``              positive float32 mesh (``FrequencyDiscretization``)
* ``cutoffs``           float32 cutoff schedule (``CutoffDiscretization``)
* ``site_rid`` / ``inv_rid``   ``getSites()`` / ``getInvertedSites()`` rids
  (``src/Lattice.hpp:481-494``), used by the bundled vertex gathers
* ``ovl_off/ovl_rid1/ovl_rid2``  overlap lists as CSR (``src/Lattice.hpp:470-474``)
* ``range_fwd/range_inv``  ``symmetryTransform(0, j)`` / ``(j, 0)`` for j in the
  range of the origin (self-energy flow, ``src/SU2/SU2FrgCore.cpp:151-155``)
* ``corr_rid``          ``symmetryTransform(b, j)`` for the correlation output
  (``src/SU2/SU2MeasurementCorrelation.cpp:161-176``)
* ``spin``, ``normalization`` core options (``src/SU2/SU2FrgCore.cpp:19-29``)
* ``coupling_zz``       per-rid sum of J_zz entering the initial condition
  (``src/SU2/SU2EffectiveAction.hpp:46-59``)
"""
from __future__ import annotations

from dataclasses import dataclass, fields
from typing import Dict, Optional

import numpy as np

from .lattice import Lattice, string_to_float
from .task import Task, parse_task


class CoreOptionError(ValueError):
    """Unknown core option (``Exception::Type::InitializationError`` in the reference)."""


@dataclass
class SU2Problem:
    mesh: np.ndarray
    cutoffs: np.ndarray
    L: int
    site_rid: np.ndarray
    inv_rid: np.ndarray
    ovl_off: np.ndarray
    ovl_rid1: np.ndarray
    ovl_rid2: np.ndarray
    range_fwd: np.ndarray
    range_inv: np.ndarray
    B: int
    Rb: int
    corr_rid: np.ndarray
    spin: np.float32
    normalization: np.float32
    coupling_zz_f32: np.ndarray   # (L,) float32 init values, accumulated in float like the reference
    coupling_zz_f64: np.ndarray   # (L,) float64 init values
    lattice: Optional[Lattice] = None

    @property
    def nw(self) -> int:
        return int(self.mesh.shape[0])

    @property
    def n_freq(self) -> int:
        """Number of frequency triples N_f = N_w^2 (N_w + 1) / 2."""
        n = self.nw
        return n * n * (n + 1) // 2

    @property
    def vertex_size(self) -> int:
        return self.n_freq * self.L

    @property
    def overlap_total(self) -> int:
        return int(self.ovl_off[-1])

    def initial_state(self, dtype=np.float64):
        """(v2, v4s, v4d) at the first cutoff: Gamma_s(rid) = J_zz(0->rid)/norm, the rest 0."""
        v2 = np.zeros(self.nw, dtype=dtype)
        per_rid = self.coupling_zz_f32 if dtype == np.float32 else self.coupling_zz_f64
        v4s = np.tile(per_rid.astype(dtype), self.n_freq)
        v4d = np.zeros(self.vertex_size, dtype=dtype)
        return v2, v4s, v4d

    def flops_per_evaluation(self, calls_s: int, calls_t: int, calls_u: int) -> float:
        """Algorithmic FLOP of one dGamma evaluation from integrand-call counts:
        94 L per S/U kernel call, 106 L + 4 O per T call, + Euler 2 |Gamma|."""
        L, O = self.L, self.overlap_total
        return (94.0 * L * (calls_s + calls_u) + (106.0 * L + 4.0 * O) * calls_t
                + 2.0 * 2 * self.vertex_size + 2.0 * self.nw)


def su2_problem(task: Task, resource_path: Optional[str] = None) -> SU2Problem:
    lattice, model, core_options = task.build_lattice(resource_path)
    return su2_problem_from_lattice(task.frequencies, task.cutoffs, lattice, core_options)


def su2_problem_from_lattice(mesh: np.ndarray, cutoffs: np.ndarray, lattice: Lattice,
                             core_options: Dict[str, str]) -> SU2Problem:
    spin = np.float32(0.5)
    normalization = np.float32(np.nan)
    for k, v in core_options.items():
        if k == "spin":
            spin = string_to_float(v)
        elif k == "normalization":
            normalization = string_to_float(v)
        elif k == "backend":
            continue  # selector for the GPU core, consumed by the factory
        else:
            raise CoreOptionError("Unknown spin model option '%s'." % k)
    if np.isnan(normalization):
        normalization = np.float32(np.float32(2.0) * spin)

    L = lattice.size
    zz32 = np.zeros(L, dtype=np.float32)
    zz64 = np.zeros(L, dtype=np.float64)
    for rid, strength in lattice.couplings:
        zz32[rid] = np.float32(zz32[rid] + np.float32(strength[0, 0] / normalization))
        zz64[rid] += float(strength[0, 0]) / float(normalization)

    off, rid1, rid2 = lattice.overlap_csr()
    r0 = lattice.ranges[0]
    range_fwd = np.array([lattice.sym_rid[0, j] for j in r0], dtype=np.int32)
    range_inv = np.array([lattice.sym_rid[j, 0] for j in r0], dtype=np.int32)
    B = len(lattice.basis)
    Rb = len(r0)
    corr = np.zeros(B * Rb, dtype=np.int32)
    for b in range(B):
        rb = lattice.ranges[b]
        if len(rb) != Rb:
            raise ValueError("basis sites with ranges of different size are not supported")
        for k, j in enumerate(rb):
            corr[b * Rb + k] = lattice.sym_rid[lattice.basis_ids[b], j]
    if (range_fwd < 0).any() or (range_inv < 0).any() or (corr < 0).any():
        raise ValueError("incomplete symmetry table")

    return SU2Problem(mesh=np.ascontiguousarray(mesh, dtype=np.float32),
                      cutoffs=np.ascontiguousarray(cutoffs, dtype=np.float32), L=L,
                      site_rid=lattice.site_rid.astype(np.int32), inv_rid=lattice.inverted_site_rid.astype(np.int32),
                      ovl_off=off, ovl_rid1=rid1, ovl_rid2=rid2, range_fwd=range_fwd, range_inv=range_inv,
                      B=B, Rb=Rb, corr_rid=corr, spin=np.float32(spin), normalization=np.float32(normalization),
                      coupling_zz_f32=zz32, coupling_zz_f64=zz64, lattice=lattice)


def su2_problem_from_xml(text_or_path: str, is_text: bool = False, resource_path: Optional[str] = None) -> SU2Problem:
    return su2_problem(parse_task(text_or_path, is_text=is_text), resource_path)


# ---------------------------------------------------------------------------
# TRI core (general time-reversal-invariant couplings, 16 spin components)
# ---------------------------------------------------------------------------
@dataclass
class TRIProblem:
    """Tables of one TRI flow (``src/TRI/TRIFrgCore.cpp``).  Besides the SU(2)
    tables it carries the spin permutations of every site transformation:
    ``getSites()[j].spinPermutation`` / ``getInvertedSites()[j].spinPermutation``
    (``src/TRI/TRIVertexTwoParticle.hpp:379-385``), the overlap transforms
    ``transformedX/Y/Z 1/2`` (``src/TRI/TRIFrgCore.cpp:739-1252``) and those of
    the correlation output (``src/TRI/TRIMeasurementCorrelation.cpp:291-313``)."""
    mesh: np.ndarray
    cutoffs: np.ndarray
    L: int
    site_rid: np.ndarray
    site_perm: np.ndarray       # (L, 3)
    inv_rid: np.ndarray
    inv_perm: np.ndarray        # (L, 3)
    ovl_off: np.ndarray
    ovl_rid1: np.ndarray
    ovl_rid2: np.ndarray
    ovl_perm1: np.ndarray       # (O, 3)
    ovl_perm2: np.ndarray       # (O, 3)
    range_fwd: np.ndarray
    range_inv: np.ndarray
    B: int
    Rb: int
    corr_rid: np.ndarray
    corr_perm: np.ndarray       # (B*Rb, 3)
    normalization: np.float32
    coupling_f32: np.ndarray    # (L, 3, 3) initial spin-spin vertex per rid, float32 as the reference adds it
    coupling_f64: np.ndarray    # (L, 3, 3)
    lattice: Optional[Lattice] = None

    components = 16

    @property
    def nw(self) -> int:
        return int(self.mesh.shape[0])

    @property
    def n_freq(self) -> int:
        n = self.nw
        return n * n * (n + 1) // 2

    @property
    def vertex_size(self) -> int:
        return self.n_freq * 16 * self.L

    @property
    def overlap_total(self) -> int:
        return int(self.ovl_off[-1])

    def initial_state(self, dtype=np.float64):
        """(v2, v4) at the first cutoff: Gamma^{s1 s2}(rid) = J_{s1 s2}(0->rid) / (4 norm)
        for spin components s1, s2 < 3, all frequencies; zero elsewhere
        (``src/TRI/TRIEffectiveAction.hpp:38-64``)."""
        v2 = np.zeros(self.nw, dtype=dtype)
        src = self.coupling_f32 if dtype == np.float32 else self.coupling_f64
        block = np.zeros((16, self.L), dtype=dtype)
        for s1 in range(3):
            for s2 in range(3):
                block[4 * s1 + s2] = src[:, s1, s2]
        v4 = np.tile(block.ravel(), self.n_freq)
        return v2, v4

    def flops_per_evaluation(self, calls_s: int, calls_t: int, calls_u: int) -> float:
        """Algorithmic FLOP of one dGamma evaluation from integrand-call counts.
        Per integrand call: 4 bundled gathers of 16 L values (2 FLOP per support
        x 4 supports: 512 L), the channel algebra (S, U: 2 x 256 products, 1024 L;
        T: chalice + inverse chalice 2 x 2 x 256 L products, 2048 L, their 4 local
        vertices 4 x 16 x 8 FLOP, and the lattice bubble 128 O products, 256 O:
        the 128 statements `RPA += 2 * a * b` per overlap entry of
        src/TRI/TRIFrgCore.cpp:739-1252, one FMA each with the 2 folded, as the
        SU(2) and XYZ bubbles are counted).  Euler 2 |Gamma|."""
        L, O = self.L, self.overlap_total
        c_su = 512.0 * L + 1024.0 * L
        c_t = 512.0 * L + 2048.0 * L + 512.0 + 256.0 * O
        return c_su * (calls_s + calls_u) + c_t * calls_t + 2.0 * self.vertex_size + 2.0 * self.nw


def tri_problem_from_lattice(mesh: np.ndarray, cutoffs: np.ndarray, lattice: Lattice,
                             core_options: Dict[str, str]) -> TRIProblem:
    normalization = np.float32(np.nan)
    for k, v in core_options.items():
        if k == "normalization":
            normalization = string_to_float(v)
        elif k == "backend":
            continue
        else:
            raise CoreOptionError("Unknown spin model option '%s'." % k)
    if np.isnan(normalization):
        normalization = np.float32(1.0)
    L = lattice.size
    c32 = np.zeros((L, 3, 3), dtype=np.float32)
    c64 = np.zeros((L, 3, 3), dtype=np.float64)
    for rid, strength in lattice.couplings:
        for s1 in range(3):
            for s2 in range(3):
                q = np.float32(np.float32(0.25) * np.float32(strength[s1, s2]))
                c32[rid, s1, s2] = np.float32(c32[rid, s1, s2] + np.float32(q / normalization))
                c64[rid, s1, s2] += 0.25 * float(strength[s1, s2]) / float(normalization)

    off, rid1, rid2 = lattice.overlap_csr()
    perm1 = np.concatenate([np.asarray(ov[2], dtype=np.int32).reshape(-1, 3) for ov in lattice.overlaps])
    perm2 = np.concatenate([np.asarray(ov[3], dtype=np.int32).reshape(-1, 3) for ov in lattice.overlaps])
    r0 = lattice.ranges[0]
    range_fwd = np.array([lattice.sym_rid[0, j] for j in r0], dtype=np.int32)
    range_inv = np.array([lattice.sym_rid[j, 0] for j in r0], dtype=np.int32)
    B = len(lattice.basis)
    Rb = len(r0)
    corr = np.zeros(B * Rb, dtype=np.int32)
    cperm = np.zeros((B * Rb, 3), dtype=np.int32)
    for b in range(B):
        rb = lattice.ranges[b]
        if len(rb) != Rb:
            raise ValueError("basis sites with ranges of different size are not supported")
        for k, j in enumerate(rb):
            corr[b * Rb + k] = lattice.sym_rid[lattice.basis_ids[b], j]
            cperm[b * Rb + k] = lattice.sym_perm[lattice.basis_ids[b], j]
    if (range_fwd < 0).any() or (range_inv < 0).any() or (corr < 0).any():
        raise ValueError("incomplete symmetry table")
    if tuple(lattice.sym_perm[0, 0]) != (0, 1, 2):
        raise ValueError("the origin pair must map to itself without spin permutation")
    return TRIProblem(mesh=np.ascontiguousarray(mesh, dtype=np.float32),
                      cutoffs=np.ascontiguousarray(cutoffs, dtype=np.float32), L=L,
                      site_rid=lattice.sym_rid[0, :L].astype(np.int32),
                      site_perm=np.ascontiguousarray(lattice.sym_perm[0, :L], dtype=np.int32),
                      inv_rid=lattice.sym_rid[:L, 0].astype(np.int32),
                      inv_perm=np.ascontiguousarray(lattice.sym_perm[:L, 0], dtype=np.int32),
                      ovl_off=off, ovl_rid1=rid1, ovl_rid2=rid2,
                      ovl_perm1=np.ascontiguousarray(perm1), ovl_perm2=np.ascontiguousarray(perm2),
                      range_fwd=range_fwd, range_inv=range_inv, B=B, Rb=Rb, corr_rid=corr,
                      corr_perm=np.ascontiguousarray(cperm), normalization=np.float32(normalization),
                      coupling_f32=c32, coupling_f64=c64, lattice=lattice)


def tri_problem(task: Task, resource_path: Optional[str] = None) -> TRIProblem:
    lattice, model, core_options = task.build_lattice(resource_path)
    return tri_problem_from_lattice(task.frequencies, task.cutoffs, lattice, core_options)


# ---------------------------------------------------------------------------
# XYZ core (diagonal couplings, 4 spin components X, Y, Z, D)
# ---------------------------------------------------------------------------
@dataclass
class XYZProblem(TRIProblem):
    """Tables of one XYZ flow (``src/XYZ/XYZFrgCore.cpp``): the lattice tables of
    the TRI core (site and overlap spin permutations act on the diagonal spin
    components X, Y, Z; the density D is permutation invariant).  The state is
    four planes X, Y, Z, D of N_f x L values (``_dataXX`` ... ``_dataDD`` of
    ``src/XYZ/XYZVertexTwoParticle.hpp:63-77``)."""

    components = 4

    @property
    def vertex_size(self) -> int:
        return self.n_freq * 4 * self.L

    def initial_state(self, dtype=np.float64):
        """(v2, v4) at the first cutoff: Gamma_c(rid) = J_cc(0->rid) / (4 norm) for
        c = X, Y, Z at all frequencies, D = 0 (``src/XYZ/XYZEffectiveAction.hpp:37-62``;
        off-diagonal couplings are ignored, as by the reference)."""
        v2 = np.zeros(self.nw, dtype=dtype)
        src = self.coupling_f32 if dtype == np.float32 else self.coupling_f64
        planes = np.zeros((4, self.n_freq, self.L), dtype=dtype)
        for c in range(3):
            planes[c] = src[:, c, c][None, :]
        return v2, planes.ravel()

    def flops_per_evaluation(self, calls_s: int, calls_t: int, calls_u: int) -> float:
        """Algorithmic FLOP of one dGamma evaluation from integrand-call counts.
        Per integrand call: 4 bundled gathers of 4 L values (2 FLOP per support
        x 4 supports: 128 L); S, U: 2 pairs x 16 products (64 L); T: chalice and
        inverse chalice 2 pairs x 2 x 16 scalar products (128 L), their 4 local
        vertices of 4 components (4 x 4 x 8 FLOP), and the lattice bubble of one
        pair (4 products per overlap entry, 8 O) scaled into the result (8 L).
        Euler 2 |Gamma|."""
        L, O = self.L, self.overlap_total
        c_su = 128.0 * L + 64.0 * L
        c_t = 128.0 * L + 128.0 * L + 128.0 + 8.0 * O + 8.0 * L
        return c_su * (calls_s + calls_u) + c_t * calls_t + 2.0 * self.vertex_size + 2.0 * self.nw


def xyz_problem_from_lattice(mesh: np.ndarray, cutoffs: np.ndarray, lattice: Lattice,
                             core_options: Dict[str, str]) -> XYZProblem:
    """XYZ tables; core options as the reference's XYZFrgCore constructor
    (``src/XYZ/XYZFrgCore.cpp:17-29``: ``normalization``, default 1)."""
    t = tri_problem_from_lattice(mesh, cutoffs, lattice, core_options)
    return XYZProblem(**{f.name: getattr(t, f.name) for f in fields(TRIProblem)})


def xyz_problem(task: Task, resource_path: Optional[str] = None) -> XYZProblem:
    lattice, model, core_options = task.build_lattice(resource_path)
    return xyz_problem_from_lattice(task.frequencies, task.cutoffs, lattice, core_options)
