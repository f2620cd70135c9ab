"""The SDF model API, mirroring the reference's constructors.

Same names, argument order and meaning as the Fortran API so scene builders read alike:
primitives `sphere, box, torus, cylinder, triprism, segment, capsule, cone, egg, plane`
(src/sdfs/sdfs.f90:158-492), `model(array, func, kopt)` (src/sdfs/sdf_base.f90:100-143), the
operators `union, SmoothUnion, subtraction, intersection` (src/sdfs/sdfModifiers.f90:412-475),
the modifiers `extrude, revolution, onion, elongate, twist, bend`
(src/sdfs/sdfModifiers.f90:125-271), the transform helpers `identity, translate, rotate_x,
rotate_y, rotate_z, rotmat, rotationAlign, skewSymm` (src/sdfs/sdfHelpers.f90:23-182), `invert`
(src/mat_class.f90:154-207) and `mono` (src/opticalProps/opticalProperties.f90:107-125).

Objects here are descriptions only: `flatten()` turns a top-level SDF into the postfix nodes of
include/smcrt.h, and evaluation happens on the GPU (signedmcrt_amd.engine.Engine.evaluate).

Matrices follow the reference's row-vector convention: a 4x4 `T` indexed T[i, j] like Fortran
T(i+1, j+1), applied as q_j = sum_i [x, y, z, 1]_i T[i, j] (src/vector_class.f90:292-304).
Not mirrored: `displacement` (needs an arbitrary Fortran function pointer) and `repeat` (an
`error stop` in the reference, src/sdfs/sdfModifiers.f90:394-410).
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np

from . import _abi as A


# ---------------------------------------------------------------------------------------------
# optical properties
# ---------------------------------------------------------------------------------------------
class mono:
    """Single-wavelength optical properties: kappa = mus + mua, albedo = mus/kappa (1 when
    mua < 1e-9), g2 = hgg**2 (src/opticalProps/opticalProperties.f90:107-125)."""

    def __init__(self, mus: float, mua: float, hgg: float, n: float):
        self.mus, self.mua, self.hgg, self.n = float(mus), float(mua), float(hgg), float(n)

    @property
    def kappa(self) -> float:
        return self.mus + self.mua

    @property
    def albedo(self) -> float:
        return 1.0 if self.mua < 1e-9 else self.mus / self.kappa

    @property
    def g2(self) -> float:
        return self.hgg ** 2

    def __repr__(self) -> str:
        return f"mono(mus={self.mus}, mua={self.mua}, hgg={self.hgg}, n={self.n})"


# ---------------------------------------------------------------------------------------------
# 4x4 helpers (src/sdfs/sdfHelpers.f90, src/mat_class.f90)
# ---------------------------------------------------------------------------------------------
def identity() -> np.ndarray:
    return np.eye(4)


def translate(o: Sequence[float]) -> np.ndarray:
    """translate(o): the offset sits in row 4 (src/sdfs/sdfHelpers.f90:170-182)."""
    t = np.eye(4)
    t[3, 0:3] = np.asarray(o, dtype=float)
    return t


def _cols(*cols) -> np.ndarray:
    return np.array(cols, dtype=float).T


def rotate_x(angle: float) -> np.ndarray:
    a = math.radians(angle)
    c, s = math.cos(a), math.sin(a)
    return _cols([1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1])


def rotate_y(angle: float) -> np.ndarray:
    a = math.radians(angle)
    c, s = math.cos(a), math.sin(a)
    return _cols([c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1])


def rotate_z(angle: float) -> np.ndarray:
    a = math.radians(angle)
    c, s = math.cos(a), math.sin(a)
    return _cols([c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1])


def rotmat(axis: Sequence[float], angle: float) -> np.ndarray:
    ax = np.asarray(axis, dtype=float)
    x, y, z = ax / np.linalg.norm(ax)
    a = math.radians(angle)
    s, c = math.sin(a), math.cos(a)
    oc = 1.0 - c
    return _cols([oc * x * x + c, oc * x * y - z * s, oc * z * x + y * s, 0.0],
                 [oc * x * y + z * s, oc * y * y + c, oc * y * z - x * s, 0.0],
                 [oc * z * x - y * s, oc * y * z + x * s, oc * z * z + c, 0.0],
                 [0.0, 0.0, 0.0, 1.0])


def skewSymm(a: Sequence[float]) -> np.ndarray:
    x, y, z = (float(v) for v in a)
    return _cols([0, -z, y, 0], [z, 0, -x, 0], [-y, x, 0, 0], [0, 0, 0, 0])


def rotationAlign(a: Sequence[float], b: Sequence[float]) -> np.ndarray:
    """Rotation taking unit vector a onto unit vector b (src/sdfs/sdfHelpers.f90:119-142)."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    v = np.cross(a, b)
    c = float(a @ b)
    k = 1.0 / (1.0 + c)
    vx = skewSymm(v)
    return np.eye(4) + vx + (vx @ vx) * k


def invert(m: np.ndarray) -> np.ndarray:
    return np.linalg.inv(np.asarray(m, dtype=float))


def _xform12(t: np.ndarray) -> List[float]:
    t = np.asarray(t, dtype=float)
    if t.shape != (4, 4):
        raise ValueError("transform must be 4x4")
    return [float(x) for x in t[:, 0:3].reshape(-1)]


# ---------------------------------------------------------------------------------------------
# SDF objects
# ---------------------------------------------------------------------------------------------
class SDF:
    """Base of every SDF description (the reference's `sdf_base`)."""

    optProps: mono
    layer: int
    transform: np.ndarray

    def flatten(self, nodes: List[A.Node], verts: List[float]) -> None:
        """Append this SDF's postfix nodes (and any polygon vertices, x y pairs) to the lists."""
        raise NotImplementedError


class _Prim(SDF):
    kind = 0

    def __init__(self, params: Sequence[float], optProp: mono, layer: int, transform=None):
        if not isinstance(optProp, mono):
            raise TypeError("optProp must be a mono(...)")
        self.params = [float(p) for p in params]
        self.optProps = optProp
        self.layer = int(layer)
        self.transform = identity() if transform is None else np.asarray(transform, dtype=float)

    def flatten(self, nodes: List[A.Node], verts: List[float]) -> None:
        n = A.Node()
        n.op, n.kind, n.k = A.OP_PRIM, self.kind, 0.0
        for i, v in enumerate(self.params):
            n.p[i] = v
        for i, v in enumerate(_xform12(self.transform)):
            n.t[i] = v
        nodes.append(n)


def _v3(a) -> List[float]:
    a = [float(x) for x in a]
    if len(a) != 3:
        raise ValueError("expected a 3-vector")
    return a


class sphere(_Prim):
    kind = A.SPHERE

    def __init__(self, radius, optProp, layer, transform=None):
        super().__init__([radius], optProp, layer, transform)


class box(_Prim):
    """box(lengths, ...): full edge lengths; halved like the reference constructor."""
    kind = A.BOX

    def __init__(self, lengths, optProp, layer, transform=None):
        super().__init__([0.5 * v for v in _v3(lengths)], optProp, layer, transform)


class torus(_Prim):
    kind = A.TORUS

    def __init__(self, oradius, iradius, optProp, layer, transform=None):
        super().__init__([oradius, iradius], optProp, layer, transform)


class cylinder(_Prim):
    kind = A.CYLINDER

    def __init__(self, a, b, radius, optProp, layer, transform=None):
        super().__init__(_v3(a) + _v3(b) + [radius], optProp, layer, transform)


class triprism(_Prim):
    kind = A.TRIPRISM

    def __init__(self, h1, h2, optProp, layer, transform=None):
        super().__init__([h1, h2], optProp, layer, transform)


class segment(_Prim):
    kind = A.SEGMENT

    def __init__(self, a, b, optProp, layer, transform=None):
        super().__init__(_v3(a) + _v3(b), optProp, layer, transform)


class capsule(_Prim):
    kind = A.CAPSULE

    def __init__(self, a, b, r, optProp, layer, transform=None):
        super().__init__(_v3(a) + _v3(b) + [r], optProp, layer, transform)


class cone(_Prim):
    kind = A.CONE

    def __init__(self, a, b, ra, rb, optProp, layer, transform=None):
        super().__init__(_v3(a) + _v3(b) + [ra, rb], optProp, layer, transform)


class egg(_Prim):
    kind = A.EGG

    def __init__(self, r1, r2, h, optProp, layer, transform=None):
        super().__init__([r1, r2, h], optProp, layer, transform)


class plane(_Prim):
    kind = A.PLANE

    def __init__(self, a, optProp, layer, transform=None):
        super().__init__(_v3(a), optProp, layer, transform)


class polygon(_Prim):
    """Closed polygon in the xy plane, `vertices` = [(x, y), ...] (simple; either orientation):
    its signed 2-D distance, z ignored — `extrude(polygon(...), h)` makes a prism of half-height
    h.  No reference counterpart (the reference's `segment` is a 0.1-thick tube, not a region);
    the library evaluates it through a per-polygon uniform-grid accelerator."""
    kind = A.POLYGON

    def __init__(self, vertices, optProp, layer, transform=None):
        v = np.asarray(vertices, dtype=float)
        if v.ndim != 2 or v.shape[1] != 2 or len(v) < 3:
            raise ValueError("polygon needs at least 3 (x, y) vertices")
        self.vertices = v
        super().__init__([0.0, float(len(v))], optProp, layer, transform)

    def flatten(self, nodes: List[A.Node], verts: List[float]) -> None:
        self.params[0] = float(len(verts) // 2)
        verts.extend(float(x) for x in self.vertices.reshape(-1))
        super().flatten(nodes, verts)


# ---- operators --------------------------------------------------------------------------------
class _Op:
    def __init__(self, kind: int, name: str, fn):
        self.kind, self.name, self._fn = kind, name, fn

    def __call__(self, d1: float, d2: float, k: float = 0.0) -> float:
        return self._fn(d1, d2, k)

    def __repr__(self) -> str:
        return self.name


def _smooth(d1, d2, k):
    h = max(k - abs(d1 - d2), 0.0) / k
    return min(d1, d2) - h * h * h * k * (1.0 / 6.0)


union = _Op(A.UNION, "union", lambda d1, d2, k: min(d1, d2))
SmoothUnion = _Op(A.SMOOTH_UNION, "SmoothUnion", _smooth)
subtraction = _Op(A.SUBTRACTION, "subtraction", lambda d1, d2, k: max(-d1, d2))
intersection = _Op(A.INTERSECTION, "intersection", lambda d1, d2, k: max(d1, d2))


class model(SDF):
    """Fold `func(d, d_i, k)` over the children left to right (src/sdfs/sdf_base.f90:145-160).
    Optical properties and layer come from the first child (:122-141)."""

    def __init__(self, array: Sequence[SDF], func: _Op, kopt: float = 0.0):
        if not array:
            raise ValueError("model needs at least one SDF")
        if not isinstance(func, _Op):
            raise TypeError("func must be one of union, SmoothUnion, subtraction, intersection")
        self.array = list(array)
        self.func = func
        self.k = float(kopt)
        self.optProps = self.array[0].optProps
        self.layer = self.array[0].layer
        self.transform = identity()

    def flatten(self, nodes: List[A.Node], verts: List[float]) -> None:
        self.array[0].flatten(nodes, verts)
        for child in self.array[1:]:
            child.flatten(nodes, verts)
            n = A.Node()
            n.op, n.kind, n.k = A.OP_CSG, self.func.kind, self.k
            nodes.append(n)


class _Mod(SDF):
    kind = 0

    def __init__(self, prim: SDF, params: Sequence[float]):
        self.prim = prim
        self.params = [float(p) for p in params]
        self.optProps = prim.optProps
        self.layer = prim.layer
        self.transform = identity()

    def flatten(self, nodes: List[A.Node], verts: List[float]) -> None:
        for op in (A.OP_MOD_BEGIN, None, A.OP_MOD_END):
            if op is None:
                self.prim.flatten(nodes, verts)
                continue
            n = A.Node()
            n.op, n.kind = op, self.kind
            for i, v in enumerate(self.params):
                n.p[i] = v
            nodes.append(n)


class extrude(_Mod):
    kind = A.EXTRUDE

    def __init__(self, prim, h):
        super().__init__(prim, [h])


class revolution(_Mod):
    kind = A.REVOLUTION

    def __init__(self, prim, o):
        super().__init__(prim, [o])


class onion(_Mod):
    kind = A.ONION

    def __init__(self, prim, thickness):
        super().__init__(prim, [thickness])


class elongate(_Mod):
    kind = A.ELONGATE

    def __init__(self, prim, size):
        super().__init__(prim, _v3(size))


class twist(_Mod):
    kind = A.TWIST

    def __init__(self, prim, k):
        super().__init__(prim, [k])


class bend(_Mod):
    kind = A.BEND

    def __init__(self, prim, k):
        super().__init__(prim, [k])
