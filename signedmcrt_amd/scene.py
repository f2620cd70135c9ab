"""Scene assembly: the SDF array, fluence grid and source flattened into `smcrt_scene_desc`.

`Scene` plays the part of `setup_simulation`'s output (src/setup.f90:14-54): the ordered array
of SDFs (order matters: it is the region order of the interface rule, src/inttau2.f90:131-140),
the `cart_grid` (src/grid.f90:53-93) and the bound emitter (src/photon.f90:129-157).  Materials
are the distinct `mono` objects in the array (shared objects share one material index, which is
the granularity of the adjoint's gradients).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _abi as A
from .sdf import SDF, mono


@dataclass
class cart_grid:
    """Voxel counts and HALF sizes of the fluence grid, centred on the origin."""
    nxg: int = 200
    nyg: int = 200
    nzg: int = 200
    xmax: float = 1.0
    ymax: float = 1.0
    zmax: float = 1.0

    @property
    def shape(self):
        return (self.nxg, self.nyg, self.nzg)

    @property
    def voxel_volume(self) -> float:
        return (2 * self.xmax / self.nxg) * (2 * self.ymax / self.nyg) * (2 * self.zmax / self.nzg)


_SOURCE_KINDS = {"pencil": A.SRC_PENCIL, "uniform": A.SRC_UNIFORM, "circular": A.SRC_CIRCULAR,
                 "point": A.SRC_POINT}


@dataclass
class source:
    """The emitter table of the TOML input (`[source]`, src/parse/parse_source.f90:17-201).

    name      pencil | uniform | circular | point (the four the loop's configs use)
    position  pencil/point origin; for circular the disc is centred at -position, as the
              reference's `circular` does (src/photon.f90:262-266)
    direction unit launch direction (pencil, uniform, circular)
    point1..3 uniform: pos = point1 + u*point2 + v*point3
    radius    circular disc radius
    """
    name: str = "point"
    position: Sequence[float] = (0.0, 0.0, 0.0)
    direction: Sequence[float] = (0.0, 0.0, -1.0)
    point1: Sequence[float] = (-1.0, -1.0, -1.0)
    point2: Sequence[float] = (2.0, 0.0, 0.0)
    point3: Sequence[float] = (0.0, 2.0, 0.0)
    radius: float = 0.5

    def to_c(self) -> A.Source:
        if self.name not in _SOURCE_KINDS:
            raise ValueError(f"No such source! ({self.name!r}; supported: "
                             f"{', '.join(_SOURCE_KINDS)})")
        s = A.Source()
        s.kind = _SOURCE_KINDS[self.name]
        for dst, src in ((s.pos, self.position), (s.dir, self.direction), (s.p1, self.point1),
                         (s.p2, self.point2), (s.p3, self.point3)):
            v = [float(x) for x in src]
            if len(v) != 3:
                raise ValueError("source vectors must have 3 components")
            for i in range(3):
                dst[i] = v[i]
        s.radius = float(self.radius)
        return s


class Scene:
    """A flattened scene plus the ctypes buffers backing its descriptor."""

    def __init__(self, array: Sequence[SDF], grid: cart_grid, src: source, detectors: Sequence = ()):
        if not array:
            raise ValueError("scene needs at least one SDF")
        self.array = list(array)
        self.grid = grid
        self.source = src
        self.detectors = list(detectors)
        if len(self.detectors) > A.MAX_DETECTORS:
            raise ValueError(f"at most {A.MAX_DETECTORS} detectors")
        self.materials: List[mono] = []
        nodes: List[A.Node] = []
        regions: List[A.Region] = []
        verts: List[float] = []
        for sdf in self.array:
            m = self._material_index(sdf.optProps)
            first = len(nodes)
            sdf.flatten(nodes, verts)
            r = A.Region()
            r.first_node, r.n_nodes, r.material, r.layer = first, len(nodes) - first, m, sdf.layer
            regions.append(r)
        self._nodes = (A.Node * len(nodes))(*nodes)
        self._regions = (A.Region * len(regions))(*regions)
        self._mats = (A.Material * max(1, len(self.materials)))()
        self._fill_materials()
        d = A.SceneDesc()
        d.abi_version = A.ABI_VERSION
        d.n_nodes, d.n_regions, d.n_materials = len(nodes), len(regions), len(self.materials)
        d.nodes = C.cast(self._nodes, C.POINTER(A.Node))
        d.regions = C.cast(self._regions, C.POINTER(A.Region))
        d.materials = C.cast(self._mats, C.POINTER(A.Material))
        g = A.Grid()
        g.nx, g.ny, g.nz = grid.nxg, grid.nyg, grid.nzg
        g.xmax, g.ymax, g.zmax = grid.xmax, grid.ymax, grid.zmax
        d.grid = g
        d.source = src.to_c()
        self._dets = (A.Detector * max(1, len(self.detectors)))(*[x.to_c() for x in self.detectors])
        d.n_detectors = len(self.detectors)
        d.detectors = C.cast(self._dets, C.POINTER(A.Detector))
        self._verts = (C.c_double * max(2, len(verts)))(*verts)
        d.n_poly_vertices = len(verts) // 2
        d.poly_vertices = C.cast(self._verts, C.POINTER(C.c_double))
        self.desc = d

    def _material_index(self, m: mono) -> int:
        for i, x in enumerate(self.materials):
            if x is m:
                return i
        self.materials.append(m)
        return len(self.materials) - 1

    def _fill_materials(self) -> None:
        for i, m in enumerate(self.materials):
            self._mats[i].mus, self._mats[i].mua = m.mus, m.mua
            self._mats[i].hgg, self._mats[i].n = m.hgg, m.n

    def material_array(self, mus: Optional[np.ndarray] = None, mua: Optional[np.ndarray] = None,
                       hgg: Optional[np.ndarray] = None, n: Optional[np.ndarray] = None):
        """A ctypes material table with properties optionally replaced (FD, adjoint, batches)."""
        arr = (A.Material * len(self.materials))()
        for i, m in enumerate(self.materials):
            arr[i].mus = float(mus[i]) if mus is not None else m.mus
            arr[i].mua = float(mua[i]) if mua is not None else m.mua
            arr[i].hgg = float(hgg[i]) if hgg is not None else m.hgg
            arr[i].n = float(n[i]) if n is not None else m.n
        return arr

    @property
    def n_regions(self) -> int:
        return len(self.array)

    @property
    def n_materials(self) -> int:
        return len(self.materials)
