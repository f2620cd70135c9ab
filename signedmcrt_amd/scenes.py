"""Scene builders: the reference's `setupGeometry` registry plus the benchmark scenes.

`build(geom_name, **params)` mirrors `setup_simulation`'s dispatch on `geom_name`
(src/setup.f90:33-52).  Reference scenes kept: `scat_test` (src/setupGeometry.f90:221-247),
`scat_test2` (:249-274), `aptran` (:147-175), `sphere_scene` (:62-106).  Added scenes (the
reference's registry has none of these): `slab`, `refracting_sphere`, `skin_vessels`,
`polygon_regions`.
Each builder returns (array, grid, source, nphotons) with the reference's defaults where the
reference fixes them and the benchmark's sizes otherwise.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

from .scene import Scene, cart_grid, source
from .sdf import (SDF, SmoothUnion, box, cylinder, extrude, invert, model, mono, polygon, sphere,
                  translate)


def scat_test(tau: float = 10.0) -> Tuple[List[SDF], cart_grid, source, int]:
    """Point source in a sphere of optical radius tau, mua = 0, g = 0 (the N_scatt KAT,
    test/end_to_end/test_scat.f90:23-41 with res/scat_test.toml)."""
    arr = [sphere(1.0, mono(tau, 0.0, 0.0, 1.0), 1),
           box((2.0, 2.0, 2.0), mono(0.0, 0.0, 0.0, 1.0), 2)]
    return arr, cart_grid(200, 200, 200, 1.0, 1.0, 1.0), source("point", (0, 0, 0)), 100000


def scat_test2(tau: float = 10.0, hgg: float = 0.9) -> Tuple[List[SDF], cart_grid, source, int]:
    """Pencil beam along +z into a near-infinite medium (the scatter-moment KAT,
    test/end_to_end/test_scat.f90:43-86 with res/scat_test2.toml)."""
    arr = [box((200.0, 200.0, 200.0), mono(tau, 1e-17, hgg, 1.0), 2)]
    return (arr, cart_grid(200, 200, 200, 100.0, 100.0, 100.0),
            source("pencil", (0, 0, 0), (0, 0, 1)), 10000000)


def aptran() -> Tuple[List[SDF], cart_grid, source, int]:
    """Refracting sphere test case of src/setupGeometry.f90:147-175 with res/aptran.toml's source
    (its vector `direction` is honoured here; see signedmcrt_amd.config)."""
    clear = mono(0.0, 1e-17, 0.0, 1.0)
    arr = [sphere(0.5, mono(0.0, 1e-17, 0.0, 1.33), 1, transform=invert(translate((0, 0, 0)))),
           box((2.0, 2.0, 2.0), clear, 2),
           box((2.01, 2.01, 2.01), mono(0.0, 10000000.0, 0.0, 1.0), 3)]
    src = source("uniform", direction=(0, 0, -1), point1=(-0.25, 0.0, 0.99999),
                 point2=(0.5, 0.0, 0.0), point3=(0.0, 0.0, 0.0))
    return arr, cart_grid(201, 201, 201, 1.0, 1.0, 1.0), src, 10000000


def sphere_scene(num_spheres: int = 10, seed: int = 123456789) -> Tuple[List[SDF], cart_grid, source, int]:
    """Random non-scattering spheres in a unit box (src/setupGeometry.f90:62-106).  The sphere
    placement draws from numpy's generator seeded by `seed` (the reference uses its global
    compiler RNG), so the exact layout differs from a reference run."""
    rng = np.random.default_rng(seed)
    inside = mono(0.0, 0.0, 0.9, 1.37)
    arr: List[SDF] = []
    for i in range(num_spheres):
        r = rng.uniform(0.001, 0.25)
        pos = [rng.uniform(-1.0 + r, 1.0 - r) for _ in range(3)]
        arr.append(sphere(r, inside, i + 1, transform=invert(translate(pos))))
    arr.append(box((2.0, 2.0, 2.0), mono(1e-17, 1e-17, 0.0, 1.0), num_spheres + 1))
    src = source("uniform", direction=(0, 0, -1), point1=(-1.0, -1.0, 0.9999999),
                 point2=(2.0, 0.0, 0.0), point3=(0.0, 2.0, 0.0))
    return arr, cart_grid(200, 200, 200, 1.0, 1.0, 1.0), src, 1000000


# ---- benchmark scenes (not in the reference registry) ----------------------------------------

def slab(mus: float = 100.0, mua: float = 0.1, hgg: float = 0.9, n: float = 1.37,
         thickness: float = 0.2, nxyz: int = 100) -> Tuple[List[SDF], cart_grid, source, int]:
    """Single slab 1 x 1 x `thickness` cm in air, normal-incidence pencil beam at the centre.
    The grid spans the slab laterally (+-0.5 cm) and +-0.25 cm in z, so it holds the source at
    z = 0.2 cm; packets leaving the grid end there (src/inttau2.f90:256-259)."""
    arr = [box((1.0, 1.0, thickness), mono(mus, mua, hgg, n), 1),
           box((2.0, 2.0, 2.0), mono(0.0, 0.0, 0.0, 1.0), 2)]
    src = source("pencil", (0.0, 0.0, 0.2), (0.0, 0.0, -1.0))
    return arr, cart_grid(nxyz, nxyz, nxyz, 0.5, 0.5, 0.25), src, 1000000


def refracting_sphere(nxyz: int = 200) -> Tuple[List[SDF], cart_grid, source, int]:
    """Clear sphere r = 0.5 cm (n = 1.5, mua = 0.01/cm) inside a 2 cm cube of scattering medium
    (mus = 50, mua = 0.05, g = 0.8, n = 1.33) in air; top-hat (circular) beam of radius 0.3 cm
    pointing -z.  The reference's circular source centres its disc at -position
    (src/photon.f90:262-266), so position (0, 0, -0.999) puts the disc at z = +0.999, just
    inside the cube top.  The grid is the cube (200^3 voxels of 0.01 cm)."""
    arr = [sphere(0.5, mono(0.0, 0.01, 0.0, 1.5), 1),
           box((2.0, 2.0, 2.0), mono(50.0, 0.05, 0.8, 1.33), 2),
           box((4.0, 4.0, 4.0), mono(0.0, 0.0, 0.0, 1.0), 3)]
    src = source("circular", position=(0.0, 0.0, -0.999), direction=(0.0, 0.0, -1.0), radius=0.3)
    return arr, cart_grid(nxyz, nxyz, nxyz, 1.0, 1.0, 1.0), src, 10000000


def skin_vessels(seed: int = 123456789, nxyz: int = 256, k: float = 0.01
                 ) -> Tuple[List[SDF], cart_grid, source, int]:
    """Five stacked tissue layers plus eight blood vessels merged by SmoothUnion, in air.

    Layer thicknesses (0.01-0.5 cm), mua (0.02-5/cm), mus (20-300/cm), g (0.8-0.95) and
    n (1.37-1.45) are drawn from numpy's generator seeded with `seed`; vessels are cylinders
    along x with r in 0.01-0.05 cm, blood mua = 20/cm, mus = 150/cm, g = 0.95, n = 1.40.
    Region order: vessels first (so a point in a vessel resolves to blood under the
    first-negative rule, src/inttau2.f90:131-140), then the layers top to bottom, then air.
    Pencil beam at the centre pointing -z, launched 0.05 cm above the skin."""
    rng = np.random.default_rng(seed)
    thick = rng.uniform(0.01, 0.5, 5)
    mua = np.exp(rng.uniform(np.log(0.02), np.log(5.0), 5))
    mus = rng.uniform(20.0, 300.0, 5)
    g = rng.uniform(0.8, 0.95, 5)
    n = rng.uniform(1.37, 1.45, 5)
    total = float(thick.sum())
    half_xy = 1.0
    top = total / 2.0
    layers: List[SDF] = []
    z = top
    for i in range(5):
        zc = z - thick[i] / 2.0
        layers.append(box((2 * half_xy, 2 * half_xy, float(thick[i])),
                          mono(float(mus[i]), float(mua[i]), float(g[i]), float(n[i])), i + 2,
                          transform=invert(translate((0.0, 0.0, zc)))))
        z -= thick[i]
    blood = mono(150.0, 20.0, 0.95, 1.40)
    cyls: List[SDF] = []
    for v in range(8):
        r = float(rng.uniform(0.01, 0.05))
        depth = float(rng.uniform(0.05, min(total, 1.0) - 0.05))
        y = float(rng.uniform(-0.8, 0.8))
        zc = top - depth
        cyls.append(cylinder((-half_xy, y, zc), (half_xy, y, zc), r, blood, 1))
    vessels = model(cyls, SmoothUnion, k)
    world = box((4.0, 4.0, 4.0 + total), mono(0.0, 0.0, 0.0, 1.0), 7)
    arr = [vessels] + layers + [world]
    zmax = top + 0.1
    src = source("pencil", (0.0, 0.0, top + 0.05), (0.0, 0.0, -1.0))
    return arr, cart_grid(nxyz, nxyz, nxyz, half_xy, half_xy, zmax), src, 100000000


def polygon_regions(seed: int = 123456789, n_side: int = 4, n_edges: int = 125, depth: float = 0.5,
                    half: float = 2.0, nxy: int = 400, nz: int = 50, mua=None
                    ) -> Tuple[List[SDF], cart_grid, source, int]:
    """n_side^2 closed polygons of n_edges vertices (2000 edges for 4 x 4 x 125), one per tile of
    an n_side x n_side tiling of [-half, half]^2, extruded to `depth`, inside a slab of background
    tissue of the same depth, in air.  Each polygon is a star-shaped perturbed circle
    r(t) = 0.38 tile (1 + sum_m a_m cos(m t + phase_m)), m = 2, 3, 5, a_m in 0.02-0.08.
    Polygon k: mu_s 50/cm, g 0.85, n 1.4, mu_a drawn log-uniform 0.05-2/cm (or mua[k]); the
    background mu_a 0.2/cm, same mu_s, g, n.  Region order: polygons, background, air.  Uniform
    source over the top face pointing -z; the grid is the slab (nxy x nxy x nz voxels)."""
    rng = np.random.default_rng(seed)
    tile = 2.0 * half / n_side
    arr: List[SDF] = []
    k = 0
    for iy in range(n_side):
        for ix in range(n_side):
            cx, cy = -half + (ix + 0.5) * tile, -half + (iy + 0.5) * tile
            amps = rng.uniform(0.02, 0.08, 3)
            phases = rng.uniform(0.0, 2.0 * np.pi, 3)
            th = 2.0 * np.pi * np.arange(n_edges) / n_edges
            r = 0.38 * tile * (1.0 + sum(a * np.cos(m * th + ph) for a, m, ph in zip(amps, (2, 3, 5), phases)))
            verts = np.stack([cx + r * np.cos(th), cy + r * np.sin(th)], axis=1)
            mu = float(mua[k]) if mua is not None else float(np.exp(rng.uniform(np.log(0.05), np.log(2.0))))
            arr.append(extrude(polygon(verts, mono(50.0, mu, 0.85, 1.4), k + 1), depth / 2.0))
            k += 1
    arr.append(box((2 * half, 2 * half, depth), mono(50.0, 0.2, 0.85, 1.4), k + 1))
    arr.append(box((2 * half + 2.0, 2 * half + 2.0, depth + 2.0), mono(0.0, 0.0, 0.0, 1.0), k + 2))
    src = source("uniform", direction=(0.0, 0.0, -1.0), point1=(-half, -half, depth / 2.0 - 1e-4),
                 point2=(2 * half, 0.0, 0.0), point3=(0.0, 2 * half, 0.0))
    return arr, cart_grid(nxy, nxy, nz, half, half, depth / 2.0), src, 100000000


REGISTRY: Dict[str, callable] = {
    "scat_test": scat_test,
    "scat_test2": scat_test2,
    "aptran": aptran,
    "sphere_scene": sphere_scene,
    "slab": slab,
    "refracting_sphere": refracting_sphere,
    "skin_vessels": skin_vessels,
    "polygon_regions": polygon_regions,
}


# [geometry] table keys each registered builder takes from a TOML input: the reference's setup
# routines read `tau` (setup_scat_test, setup_scat_test2), `hgg` (setup_scat_test2) and
# `num_spheres` (setup_sphere_scene) from it (src/setupGeometry.f90); the others take none.
TOML_PARAMS: Dict[str, Tuple[str, ...]] = {
    "scat_test": ("tau",),
    "scat_test2": ("tau", "hgg"),
    "sphere_scene": ("num_spheres",),
}


def build(geom_name: str, **params) -> Tuple[Scene, int]:
    """Scene and default packet count for a registered geometry ("no such routine" otherwise,
    src/setup.f90:50-51)."""
    if geom_name not in REGISTRY:
        raise ValueError(f"no such routine: {geom_name!r}")
    arr, grid, src, n = REGISTRY[geom_name](**params)
    return Scene(arr, grid, src), n
