"""Detectors (src/detectors/detectors.f90): the reference's constructors, as descriptions that
cross the C ABI as `smcrt_detector`.

    circle_dect(pos, dir, layer, radius, nbins, maxval)        init_circle_dect
    annulus_dect(pos, dir, layer, r1, r2, nbins, maxval)       init_annulus_dect
    camera(p1, p2, p3, layer, nbins, maxval)                   init_camera

`layer` is the reference's 1-based layer (index of the SDF in the scene array, plus one).  The
hit test and binning run on the GPU at the end of every history (see include/smcrt.h);
`Engine.run` returns each detector's bins: nbins+1 counts for circle and annulus, an
(nbins+1, nbins+1) array for the camera (the reference's `data(idx, idy)`).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

from . import _abi as A


def _v3(a) -> list:
    v = [float(x) for x in a]
    if len(v) != 3:
        raise ValueError("expected a 3-vector")
    return v


@dataclass
class _Det:
    kind: int
    layer: int
    nbins: int
    maxval: float
    pos: Sequence[float] = (0.0, 0.0, 0.0)
    dir: Sequence[float] = (0.0, 0.0, -1.0)
    p2: Sequence[float] = (0.0, 0.0, 0.0)
    p3: Sequence[float] = (0.0, 0.0, 0.0)
    radius1: float = 0.0
    radius2: float = 0.0
    track_history: bool = False

    def to_c(self) -> A.Detector:
        d = A.Detector()
        d.kind, d.layer, d.nbins, d.track_history = self.kind, int(self.layer), int(self.nbins), int(self.track_history)
        for dst, src in ((d.pos, self.pos), (d.dir, self.dir), (d.p2, self.p2), (d.p3, self.p3)):
            for i, v in enumerate(_v3(src)):
                dst[i] = v
        d.radius1, d.radius2, d.maxval = float(self.radius1), float(self.radius2), float(self.maxval)
        return d

    @property
    def shape(self):
        nb = self.nbins + 1
        return (nb, nb) if self.kind == A.DET_CAMERA else (nb,)


def circle_dect(pos, dir, layer: int, radius: float, nbins: int = 100, maxval: float = 100.0,
                trackHistory: bool = False) -> _Det:
    return _Det(A.DET_CIRCLE, layer, nbins, maxval, pos=_v3(pos), dir=_v3(dir), radius1=radius,
                track_history=trackHistory)


def annulus_dect(pos, dir, layer: int, r1: float, r2: float, nbins: int = 100, maxval: float = 100.0,
                 trackHistory: bool = False) -> _Det:
    if not r2 > r1:
        raise ValueError(f"Radii are invalid: Expected radius2 ({r2}) > radius 1 ({r1})")
    return _Det(A.DET_ANNULUS, layer, nbins, maxval, pos=_v3(pos), dir=_v3(dir), radius1=r1, radius2=r2,
                track_history=trackHistory)


def camera(p1, p2, p3, layer: int, nbins: int = 100, maxval: float = 100.0,
           trackHistory: bool = False) -> _Det:
    return _Det(A.DET_CAMERA, layer, nbins, maxval, pos=_v3(p1), p2=_v3(p2), p3=_v3(p3),
                track_history=trackHistory)


def from_c(d: A.Detector) -> _Det:
    return _Det(int(d.kind), int(d.layer), int(d.nbins), float(d.maxval), pos=list(d.pos), dir=list(d.dir),
                p2=list(d.p2), p3=list(d.p3), radius1=float(d.radius1), radius2=float(d.radius2),
                track_history=bool(d.track_history))
