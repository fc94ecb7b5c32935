"""Boost-free lattice / spin-model factory.

Builds the symmetry-reduced lattice tables that the flow kernels consume, from
unit-cell and model definitions written in the reference's resource XML format
(``res/lattices.xml``, ``res/models.xml``).  The construction follows the
reference's ``LatticeModelFactory`` step for step, because the numbering of the
representative sites (rid), the spin permutations and the order of every table
are part of the on-disk formats (vertex layout, ``.obs`` site order):

* neighbourhoods by breadth-first range growth  -- ``src/LatticeModelFactory.cpp:458-489``
* site lookup by fractional coordinates          -- ``src/LatticeModelFactory.cpp:492-513``
* automorphism search (site2 -> origin)          -- ``src/LatticeModelFactory.cpp:516-730``
* equivalence classes, rid order, symmetry table,
  site / inverted-site buffers, overlap lists,
  per-rid couplings                              -- ``src/LatticeModelFactory.cpp:733-985``
* arithmetic-expression attribute parsing        -- ``src/lib/InputParser.hpp:21-63``

This is host-side set-up code; it runs once per flow and never touches the GPU.
"""
from __future__ import annotations

import math
import os
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

EPSILON = 0.00001
PI = 3.14159265358979323846

Site = Tuple[int, int, int, int]  # (a0, a1, a2, b)
Vec = Tuple[float, float, float]

_RES_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "res")


class LatticeError(RuntimeError):
    """Raised where the reference throws ``Exception::Type::InitializationError``."""


# ----------------------------------------------------------------------------
# expression parsing (src/lib/InputParser.hpp:21-63)
# ----------------------------------------------------------------------------
_RE_SQRT = re.compile(r"sqrt\(([\.\d]+)\)")
_RE_MULDIV = re.compile(r"([\.\d]+)([\*/])([\.\d]+)")


def string_to_double(text: str) -> float:
    """Evaluate the reference's attribute mini-language.

    ``sqrt(x)`` terms are substituted first, then the leftmost ``a*b`` / ``a/b``
    pair is folded repeatedly; every intermediate is re-printed with 20 fixed
    decimals before being parsed again, exactly as the reference's
    ``std::stringstream << std::fixed << std::setprecision(20)`` round trip
    (``src/lib/InputParser.hpp:36-58``).
    """
    s = text.strip()
    while True:
        m = _RE_SQRT.search(s)
        if not m:
            break
        s = s[: m.start()] + ("%.20f" % math.sqrt(float(m.group(1)))) + s[m.end():]
    while True:
        m = _RE_MULDIV.search(s)
        if not m:
            break
        a, b = float(m.group(1)), float(m.group(3))
        v = a * b if m.group(2) == "*" else a / b
        s = s[: m.start()] + ("%.20f" % v) + s[m.end():]
    return float(s)


def string_to_float(text: str) -> np.float32:
    """``InputParser::stringToFloat``: double evaluation, then rounding to float."""
    return np.float32(string_to_double(text))


# ----------------------------------------------------------------------------
# small double-precision geometry (src/lib/Geometry.hpp), same operation order
# ----------------------------------------------------------------------------
def _add(a: Vec, b: Vec) -> Vec:
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _sub(a: Vec, b: Vec) -> Vec:
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _scale(s: float, a: Vec) -> Vec:
    return (s * a[0], s * a[1], s * a[2])


def _dot(a: Vec, b: Vec) -> float:
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a: Vec, b: Vec) -> Vec:
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _norm(a: Vec) -> float:
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _normalize(a: Vec) -> Vec:
    n = _norm(a)
    return (a[0] / n, a[1] / n, a[2] / n)


def _acos(x: float) -> float:
    # C's acos returns NaN outside [-1, 1]; the reference relies on that silently
    return math.acos(x) if -1.0 <= x <= 1.0 else float("nan")


def _round_half_away(x: float) -> float:
    return math.copysign(math.floor(abs(x) + 0.5), x)


Mat4 = List[List[float]]


def _m4(*v: float) -> Mat4:
    return [list(v[0:4]), list(v[4:8]), list(v[8:12]), list(v[12:16])]


def _m4_identity() -> Mat4:
    return _m4(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)


def _m4_inversion() -> Mat4:
    return _m4(-1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1)


def _m4_translation(v: Vec) -> Mat4:
    return _m4(1, 0, 0, v[0], 0, 1, 0, v[1], 0, 0, 1, v[2], 0, 0, 0, 1)


def _m4_mul(a: Mat4, b: Mat4) -> Mat4:
    m = [[0.0] * 4 for _ in range(4)]
    for i in range(4):
        for j in range(4):
            acc = 0.0
            for k in range(4):
                acc += a[i][k] * b[k][j]
            m[i][j] = acc
    return m


def _m4_rotation_origin(axis: Vec, angle: float) -> Mat4:
    a = _normalize(axis)
    c, s = math.cos(angle), math.sin(angle)
    return _m4(
        a[0] * a[0] * (1 - c) + c, a[0] * a[1] * (1 - c) - a[2] * s, a[0] * a[2] * (1 - c) + a[1] * s, 0,
        a[1] * a[0] * (1 - c) + a[2] * s, a[1] * a[1] * (1 - c) + c, a[1] * a[2] * (1 - c) - a[0] * s, 0,
        a[2] * a[0] * (1 - c) - a[1] * s, a[2] * a[1] * (1 - c) + a[0] * s, a[2] * a[2] * (1 - c) + c, 0,
        0, 0, 0, 1)


def _m4_rotation(axis: Vec, point: Vec, angle: float) -> Mat4:
    return _m4_mul(_m4_mul(_m4_translation(point), _m4_rotation_origin(axis, angle)),
                   _m4_translation((-point[0], -point[1], -point[2])))


def _m4_apply(m: Mat4, v: Vec) -> Vec:
    return (m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2] + m[0][3],
            m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2] + m[1][3],
            m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2] + m[2][3])


def _m3_inverse_cols(a0: Vec, a1: Vec, a2: Vec):
    """Inverse of the matrix whose columns are a0, a1, a2 (``Mat3::inverse``)."""
    d = [[a0[0], a1[0], a2[0]], [a0[1], a1[1], a2[1]], [a0[2], a1[2], a2[2]]]
    det = (-d[0][2] * d[1][1] * d[2][0] + d[0][1] * d[1][2] * d[2][0] + d[0][2] * d[1][0] * d[2][1]
           - d[0][0] * d[1][2] * d[2][1] - d[0][1] * d[1][0] * d[2][2] + d[0][0] * d[1][1] * d[2][2])
    inv = 1 / det
    adj = [
        [-d[1][2] * d[2][1] + d[1][1] * d[2][2], d[0][2] * d[2][1] - d[0][1] * d[2][2], -d[0][2] * d[1][1] + d[0][1] * d[1][2]],
        [d[1][2] * d[2][0] - d[1][0] * d[2][2], -d[0][2] * d[2][0] + d[0][0] * d[2][2], d[0][2] * d[1][0] - d[0][0] * d[1][2]],
        [-d[1][1] * d[2][0] + d[1][0] * d[2][1], d[0][1] * d[2][0] - d[0][0] * d[2][1], -d[0][1] * d[1][0] + d[0][0] * d[1][1]],
    ]
    return [[inv * adj[i][j] for j in range(3)] for i in range(3)]


# ----------------------------------------------------------------------------
# spin permutations (src/LatticeModelFactory.cpp:369-455)
# ----------------------------------------------------------------------------
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
IDENTITY = (0, 1, 2)


def _perm_inverse(p: Tuple[int, int, int]) -> Tuple[int, int, int]:
    q = [0, 1, 2]
    for k in range(3):
        q[p[k]] = k
    return (q[0], q[1], q[2])


# ----------------------------------------------------------------------------
# unit cells and models
# ----------------------------------------------------------------------------
@dataclass
class LatticeUnitCell:
    """Bravais vectors, basis positions and nearest-neighbour bonds.

    Mirrors ``LatticeModelFactory::LatticeUnitCell`` (read at
    ``src/LatticeModelFactory.cpp:107-181``); a bond is (fromB, toB, da0, da1, da2).
    """

    lattice_vectors: List[Vec] = field(default_factory=list)
    basis_sites: List[Vec] = field(default_factory=list)
    bonds: List[Tuple[int, int, int, int, int]] = field(default_factory=list)

    @classmethod
    def from_resources(cls, name: str, bundle: Optional[str] = None) -> "LatticeUnitCell":
        for root in _iter_resource_roots(bundle):
            for uc in root:
                if uc.tag != "unitcell":
                    continue
                if uc.get("name") is None:
                    raise LatticeError("lattice unit cell without a name attribute")
                if uc.get("name") != name:
                    continue
                out = cls()
                prims = [c for c in uc if c.tag == "primitive"]
                if len(prims) != 3:
                    raise LatticeError("unit cell '%s' needs exactly three primitive vectors" % name)
                if not any(c.tag == "site" for c in uc):
                    raise LatticeError("unit cell '%s' has no lattice site" % name)
                if not any(c.tag == "bond" for c in uc):
                    raise LatticeError("unit cell '%s' has no lattice bond" % name)
                for c in uc:
                    if c.tag in ("primitive", "site"):
                        v = (string_to_double(c.get("x")), string_to_double(c.get("y")), string_to_double(c.get("z")))
                        (out.lattice_vectors if c.tag == "primitive" else out.basis_sites).append(v)
                    elif c.tag == "bond":
                        out.bonds.append((int(c.get("from")), int(c.get("to")), int(c.get("da0")),
                                          int(c.get("da1")), int(c.get("da2"))))
                return out
        raise LatticeError("no lattice unit cell named '%s'" % name)


@dataclass
class SpinInteraction:
    """A 3x3 coupling between two lattice sites, float32 like the reference."""

    frm: Site
    to: Site
    strength: np.ndarray  # (3, 3) float32

    def is_connecting_from_to(self, site_from: Site, site_to: Site) -> bool:
        return ((self.to[0] - self.frm[0]) == (site_to[0] - site_from[0])
                and (self.to[1] - self.frm[1]) == (site_to[1] - site_from[1])
                and (self.to[2] - self.frm[2]) == (site_to[2] - site_from[2])
                and self.frm[3] == site_from[3] and self.to[3] == site_to[3])

    def is_connecting_sites(self, s1: Site, s2: Site) -> int:
        if self.is_connecting_from_to(s1, s2):
            return 1
        if self.is_connecting_from_to(s2, s1):
            return -1
        return 0


_RE_SITE = re.compile(r"(-{0,1}\d+)[, ](-{0,1}\d+)[, ](-{0,1}\d+)[, ](\d+)")
_RE_COMPONENTS = re.compile(r"(-?)([xyz])([xyz])")


@dataclass
class SpinModelUnitCell:
    """Spin couplings of a model (``src/LatticeModelFactory.cpp:244-365``)."""

    interactions: List[SpinInteraction] = field(default_factory=list)
    interaction_parameters: List[str] = field(default_factory=list)

    def add(self, frm: Site, to: Site, strength: np.ndarray) -> None:
        new = SpinInteraction(frm, to, np.asarray(strength, dtype=np.float32).copy())
        for j in self.interactions:
            if j.is_connecting_sites(new.frm, new.to) != 0:
                # SpinInteraction::operator+= (src/LatticeModelFactory.cpp:208-233)
                if j.frm == new.frm and j.to == new.to:
                    j.strength = (j.strength + new.strength).astype(np.float32)
                else:
                    j.strength = (j.strength + new.strength.T).astype(np.float32)
                return
        self.interactions.append(new)

    @classmethod
    def from_resources(cls, name: str, options: Dict[str, str], bundle: Optional[str] = None) -> "SpinModelUnitCell":
        for root in _iter_resource_roots(bundle):
            for model in root:
                if model.tag != "model":
                    continue
                if model.get("name") is None:
                    raise LatticeError("spin model without a name attribute")
                if model.get("name") != name:
                    continue
                out = cls()
                if not any(c.tag == "interaction" for c in model):
                    raise LatticeError("spin model '%s' has no interaction" % name)
                for c in model:
                    if c.tag != "interaction":
                        continue
                    param, frm, to, typ = c.get("parameter"), c.get("from"), c.get("to"), c.get("type")
                    if None in (param, frm, to, typ):
                        raise LatticeError("an interaction of spin model '%s' lacks its parameter, from, to or type" % name)
                    m1, m2 = _RE_SITE.fullmatch(frm), _RE_SITE.fullmatch(to)
                    if not m1 or not m2:
                        raise LatticeError("interaction sites '%s' / '%s' of spin model '%s' cannot be parsed" % (frm, to, name))
                    if param not in options:
                        raise LatticeError("the task file sets no value for interaction parameter '%s'" % param)
                    val = string_to_float(options[param])
                    st = np.zeros((3, 3), dtype=np.float32)
                    if typ == "heisenberg":
                        st[0, 0] += val; st[1, 1] += val; st[2, 2] += val
                    elif typ == "xxyy":
                        st[0, 0] += val; st[1, 1] += val
                    elif typ == "gx":
                        st[1, 2] += val; st[2, 1] += val
                    elif typ == "gy":
                        st[2, 0] += val; st[0, 2] += val
                    elif typ == "gz":
                        st[0, 1] += val; st[1, 0] += val
                    else:
                        mc = _RE_COMPONENTS.fullmatch(typ)
                        if not mc:
                            raise LatticeError("unknown interaction type '%s' in spin model '%s'" % (typ, name))
                        sign = np.float32(-1.0 if mc.group(1) == "-" else 1.0)
                        st["xyz".index(mc.group(2)), "xyz".index(mc.group(3))] += sign * val
                    out.add(tuple(int(x) for x in m1.groups()), tuple(int(x) for x in m2.groups()), st)
                    if param not in out.interaction_parameters:
                        out.interaction_parameters.append(param)
                return out
        raise LatticeError("no spin model named '%s'" % name)


def _iter_resource_roots(bundle: Optional[str]):
    paths = [bundle] if bundle else [_RES_DIR]
    for path in paths:
        for fn in sorted(os.listdir(path)):
            if fn.endswith(".xml"):
                # the resource files hold several top-level elements; wrap them
                with open(os.path.join(path, fn), "r", encoding="utf-8") as fh:
                    text = fh.read()
                text = re.sub(r"<\?xml[^>]*\?>", "", text)
                yield ET.fromstring("<root>" + text + "</root>")


# ----------------------------------------------------------------------------
# the factory
# ----------------------------------------------------------------------------
class _Geometry:
    def __init__(self, uc: LatticeUnitCell):
        self.uc = uc
        self.inv = _m3_inverse_cols(*uc.lattice_vectors)

    def pos(self, s: Site) -> Vec:
        v = self.uc.lattice_vectors
        p = _add(_add(_scale(float(s[0]), v[0]), _scale(float(s[1]), v[1])), _scale(float(s[2]), v[2]))
        return _add(p, self.uc.basis_sites[s[3]])

    def site_at(self, p: Vec) -> Optional[Site]:
        inv = self.inv
        for b, bp in enumerate(self.uc.basis_sites):
            d = _sub(p, bp)
            n = (inv[0][0] * d[0] + inv[0][1] * d[1] + inv[0][2] * d[2],
                 inv[1][0] * d[0] + inv[1][1] * d[1] + inv[1][2] * d[2],
                 inv[2][0] * d[0] + inv[2][1] * d[1] + inv[2][2] * d[2])
            r = [_round_half_away(x) for x in n]
            if all(abs(n[k] - r[k]) < EPSILON for k in range(3)):
                return (int(r[0]), int(r[1]), int(r[2]), b)
        return None

    def neighbors(self, s: Site) -> List[Site]:
        out: List[Site] = []
        for (fb, tb, d0, d1, d2) in self.uc.bonds:
            if fb == s[3] or tb == s[3]:
                if fb == s[3]:
                    out.append((s[0] + d0, s[1] + d1, s[2] + d2, tb))
                    if fb == tb:
                        out.append((s[0] - d0, s[1] - d1, s[2] - d2, fb))
                else:
                    out.append((s[0] - d0, s[1] - d1, s[2] - d2, fb))
        return out

    def range_around(self, s: Site, rng: int) -> List[Site]:
        sites = [s]
        seen = {s}
        for _ in range(rng):
            nbs: List[Site] = []
            nbs_seen = set()
            for x in sites:
                for n in self.neighbors(x):
                    if n not in nbs_seen:
                        nbs_seen.add(n)
                        nbs.append(n)
            for n in nbs:
                if n not in seen:
                    seen.add(n)
                    sites.append(n)
        return sites


def _symmetry_reduce(g: _Geometry, model: SpinModelUnitCell, site1: Site, site2: Site,
                     reduced: bool = False) -> List[Tuple[Site, Tuple[int, int, int]]]:
    """Images f_i(site1) for all automorphisms f_i with f_i(site2) = origin.

    Restates ``symmetryReduce`` (``src/LatticeModelFactory.cpp:516-730``).
    """
    uc = g.uc
    ref: Site = (0, 0, 0, 0)
    pre_sites: List[Site] = []
    for b in range(len(uc.basis_sites)):
        if (0, 0, 0, b) not in pre_sites:
            pre_sites.append((0, 0, 0, b))
    for (fb, tb, d0, d1, d2) in uc.bonds:
        if (d0, d1, d2, tb) not in pre_sites:
            pre_sites.append((d0, d1, d2, tb))

    nbs = g.neighbors(site2)
    if len(nbs) < 3:
        raise LatticeError("lattice construction: a site has fewer than three neighbours")
    n1 = nbs[0]
    n2 = None
    p2 = g.pos(site2)
    for i in range(1, len(nbs)):
        if _norm(_cross(_sub(g.pos(n1), p2), _sub(g.pos(nbs[i]), p2))) > EPSILON:
            n2 = nbs[i]
            break
        if i == len(nbs) - 1:
            raise LatticeError("lattice construction: all neighbours of a site lie on one line")

    pref = g.pos(ref)
    valid: List[Tuple[Mat4, Tuple[int, int, int]]] = []
    ref_nbs = g.neighbors(ref)
    for c1 in ref_nbs:
        for c2 in ref_nbs:
            pc1, pc2 = g.pos(c1), g.pos(c2)
            if _norm(_cross(_sub(pc1, pref), _sub(pc2, pref))) < EPSILON:
                continue
            for inversion in (0, 1):
                if reduced and len(valid) > 0:
                    break
                t = _m4_inversion() if inversion else _m4_identity()
                t = _m4_mul(_m4_translation(_sub(pref, _m4_apply(t, p2))), t)

                a = _normalize(_sub(_m4_apply(t, g.pos(n1)), pref))
                b = _normalize(_sub(pc1, pref))
                axis = _cross(a, b)
                angle = _acos(_dot(a, b))
                if _norm(axis) < EPSILON:
                    if _dot(a, b) < 0:
                        axis = (-a[1], a[0], 0.0) if a[0] != 0 else (0.0, -a[2], a[1])
                        angle = PI
                    else:
                        axis = (1.0, 0.0, 0.0)
                        angle = 0.0
                t = _m4_mul(_m4_rotation(axis, pref, angle), t)

                c1n = _normalize(_sub(pc1, pref))
                tn2 = _sub(_m4_apply(t, g.pos(n2)), pref)
                x = _normalize(_sub(tn2, _scale(_dot(tn2, c1n), c1n)))
                c2r = _sub(pc2, pref)
                y = _normalize(_sub(c2r, _scale(_dot(c2r, c1n), c1n)))
                axis = _cross(x, y)
                angle = _acos(_dot(x, y))
                if _norm(axis) < EPSILON:
                    axis = b
                    angle = PI if _dot(x, y) < 0 else 0.0
                t = _m4_mul(_m4_rotation(axis, pref, angle), t)

                e = _norm(_sub(_m4_apply(t, p2), pref))
                if e > EPSILON:
                    raise LatticeError("symmetry construction: the transform does not carry the site onto the reference site")
                e = _norm(_cross(_normalize(_sub(_m4_apply(t, g.pos(n1)), _m4_apply(t, p2))), _normalize(_sub(pc1, pref))))
                if e > EPSILON:
                    raise LatticeError("symmetry construction: the transform does not align the first neighbour")
                e = _dot(_sub(_m4_apply(t, g.pos(n2)), pref), _cross(_sub(pc1, pref), _sub(pc2, pref)))
                if abs(e) > EPSILON:
                    raise LatticeError("symmetry construction: the second neighbour leaves the target plane")

                if any(g.site_at(_m4_apply(t, g.pos(p))) is None for p in pre_sites):
                    continue

                spin_t = None
                for perm in _PERMS:
                    ok = True
                    for inter in model.interactions:
                        ft = g.site_at(_m4_apply(t, g.pos(inter.frm))) or (0, 0, 0, 0)
                        tt = g.site_at(_m4_apply(t, g.pos(inter.to))) or (0, 0, 0, 0)
                        target = np.zeros((3, 3), dtype=np.float32)
                        for i2 in model.interactions:
                            conn = i2.is_connecting_sites(ft, tt)
                            if conn == 1:
                                target = i2.strength
                            elif conn == -1:
                                target = i2.strength.T
                        for s1 in range(3):
                            for s2 in range(3):
                                if inter.strength[s1, s2] != target[perm[s1], perm[s2]]:
                                    ok = False
                    if ok:
                        spin_t = perm
                        break
                if spin_t is None:
                    continue
                valid.append((t, spin_t))

    if not valid:
        raise LatticeError("symmetry construction: not even the identity was found valid")
    out: List[Tuple[Site, Tuple[int, int, int]]] = []
    p1 = g.pos(site1)
    for (t, perm) in valid:
        ts1 = g.site_at(_m4_apply(t, p1))
        if ts1 is None:
            raise LatticeError("symmetry construction: a valid transform maps the site outside the lattice")
        if (ts1, perm) not in out:
            out.append((ts1, perm))
    return out


@dataclass
class Lattice:
    """Symmetry-reduced lattice tables (the fields of the reference's ``Lattice``,
    ``src/Lattice.hpp:297-552``), held as numpy arrays.

    ``sym_rid[i, j]`` / ``sym_perm[i, j]`` is the transformation mapping the pair
    (site_i, site_j) to (origin, rid); ``-1`` where the pair is out of range.
    """

    size: int                       # number of representatives L
    sites: List[Site]               # all stored sites (geometry table)
    bravais: List[Vec]
    basis: List[Vec]
    basis_ids: List[int]            # site index of (0,0,0,b)
    ranges: List[List[int]]         # site indices in range of basis site b
    sym_rid: np.ndarray             # (N, N) int32
    sym_perm: np.ndarray            # (N, N, 3) int8
    overlaps: List[Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]]  # per rid: rid1, rid2, perm1, perm2
    couplings: List[Tuple[int, np.ndarray]]  # (rid, 3x3 float32) as in SpinModel::interactions

    @property
    def data_size(self) -> int:
        return len(self.sites)

    def position(self, i: int) -> Vec:
        s = self.sites[i]
        v = self.bravais
        p = _add(_add(_scale(float(s[0]), v[0]), _scale(float(s[1]), v[1])), _scale(float(s[2]), v[2]))
        return _add(p, self.basis[s[3]])

    def site_index(self, s: Site) -> int:
        return self.sites.index(s)

    def symmetry_transform(self, i1: int, i2: int) -> int:
        rid = int(self.sym_rid[i1, i2])
        if rid < 0:
            raise LatticeError("pair out of range")
        return rid

    @property
    def site_rid(self) -> np.ndarray:
        """``getSites()[j].rid`` for j < L (``src/Lattice.hpp:491-494``)."""
        return self.sym_rid[0, : self.size].copy()

    @property
    def inverted_site_rid(self) -> np.ndarray:
        """``getInvertedSites()[j].rid`` for j < L (``src/Lattice.hpp:481-484``)."""
        return self.sym_rid[: self.size, 0].copy()

    def overlap_csr(self):
        off = np.zeros(self.size + 1, dtype=np.int32)
        for r, ov in enumerate(self.overlaps):
            off[r + 1] = off[r] + len(ov[0])
        rid1 = np.concatenate([ov[0] for ov in self.overlaps]).astype(np.int32)
        rid2 = np.concatenate([ov[1] for ov in self.overlaps]).astype(np.int32)
        return off, rid1, rid2


def new_lattice_model(uc: LatticeUnitCell, model: SpinModelUnitCell, lattice_range: int) -> Lattice:
    """``LatticeModelFactory::newLatticeModel`` (``src/LatticeModelFactory.cpp:733-985``)."""
    g = _Geometry(uc)
    nb_count = len(uc.basis_sites)
    neighborhoods = [g.range_around((0, 0, 0, b), lattice_range) for b in range(nb_count)]
    nb_sets = [set(n) for n in neighborhoods]

    nb0 = neighborhoods[0]
    eq_rid = [-1] * len(nb0)
    eq_perm = [IDENTITY] * len(nb0)
    pos0 = {s: i for i, s in enumerate(nb0)}
    rid = 0
    for i in range(len(nb0)):
        if eq_rid[i] != -1:
            continue
        equiv = _symmetry_reduce(g, model, nb0[i], (0, 0, 0, 0))
        eq_rid[i] = rid
        eq_perm[i] = IDENTITY
        for (site, perm) in equiv:
            j = pos0.get(site)
            if j is None:
                raise LatticeError("symmetry image outside of the lattice range")
            if eq_rid[j] != -1:
                continue
            eq_rid[j] = rid
            eq_perm[j] = _perm_inverse(perm)
        rid += 1
    # trivial representatives first (src/LatticeModelFactory.cpp:782-793), including
    # the repeated swaps of the reference's inner loop
    nb0 = list(nb0)
    for r in range(rid):
        for i in range(len(eq_rid)):
            if eq_rid[i] == r and eq_perm[i] == IDENTITY:
                eq_rid[r], eq_rid[i] = eq_rid[i], eq_rid[r]
                eq_perm[r], eq_perm[i] = eq_perm[i], eq_perm[r]
                nb0[r], nb0[i] = nb0[i], nb0[r]
    neighborhoods[0] = nb0
    nb_sets[0] = set(nb0)
    size = rid

    sites = list(nb0)
    index = {s: i for i, s in enumerate(sites)}
    for b in range(nb_count):
        for n in neighborhoods[b]:
            if n not in index:
                index[n] = len(sites)
                sites.append(n)
    n_sites = len(sites)
    basis_ids = [index[(0, 0, 0, b)] for b in range(nb_count)]
    ranges = [[index[n] for n in neighborhoods[b]] for b in range(nb_count)]

    sym_rid = np.full((n_sites, n_sites), -1, dtype=np.int32)
    sym_perm = np.zeros((n_sites, n_sites, 3), dtype=np.int8)
    sym_perm[:, :] = IDENTITY
    for n in range(len(nb0)):
        sym_rid[0, n] = eq_rid[n]
        sym_perm[0, n] = eq_perm[n]
    for b in range(1, nb_count):
        for n in neighborhoods[b]:
            equiv = _symmetry_reduce(g, model, n, (0, 0, 0, b), True)
            if not equiv:
                raise LatticeError("lattice construction: no symmetry maps a site pair into the reduced set")
            bid, nid, eid = index[(0, 0, 0, b)], index[n], index[equiv[0][0]]
            sym_rid[bid, nid] = sym_rid[0, eid]
            p = equiv[0][1]
            sym_perm[bid, nid] = (eq_perm[eid][p[0]], eq_perm[eid][p[1]], eq_perm[eid][p[2]])
    for s1 in range(n_sites):
        a = sites[s1]
        s1pid = basis_ids[a[3]]
        nbs = nb_sets[a[3]]
        for s2 in range(n_sites):
            c = sites[s2]
            s2p = (c[0] - a[0], c[1] - a[1], c[2] - a[2], c[3])
            if s2p not in nbs:
                continue
            s2pid = index[s2p]
            sym_rid[s1, s2] = sym_rid[s1pid, s2pid]
            sym_perm[s1, s2] = sym_perm[s1pid, s2pid]

    overlaps = []
    i1 = sites[0]
    for r in range(size):
        i2 = sites[r]
        r1, r2, p1, p2 = [], [], [], []
        for j in range(n_sites):
            sj = sites[j]
            if sj not in nb_sets[i1[3]]:
                continue
            if (sj[0] - i2[0], sj[1] - i2[1], sj[2] - i2[2], sj[3]) not in nb_sets[i2[3]]:
                continue
            r1.append(sym_rid[0, j]); p1.append(sym_perm[0, j])
            r2.append(sym_rid[j, r]); p2.append(sym_perm[j, r])
        overlaps.append((np.array(r1, dtype=np.int32), np.array(r2, dtype=np.int32),
                         np.array(p1, dtype=np.int8).reshape(-1, 3), np.array(p2, dtype=np.int8).reshape(-1, 3)))

    return Lattice(size=size, sites=sites, bravais=list(uc.lattice_vectors), basis=list(uc.basis_sites),
                   basis_ids=basis_ids, ranges=ranges, sym_rid=sym_rid, sym_perm=sym_perm,
                   overlaps=overlaps, couplings=model_couplings(sites[:size], model))


def model_couplings(rep_sites: Sequence[Site], model: SpinModelUnitCell) -> List[Tuple[int, np.ndarray]]:
    """Couplings (origin -> representative rid) in the order of the reference's
    ``SpinModel::interactions`` (``src/LatticeModelFactory.cpp:957-985``)."""
    couplings = []
    for r, s in enumerate(rep_sites):
        for inter in model.interactions:
            conn = inter.is_connecting_sites((0, 0, 0, 0), s)
            if conn == 1:
                couplings.append((r, inter.strength.copy()))
            elif conn == -1:
                couplings.append((r, inter.strength.T.copy()))
    return couplings
