"""The reference's TOML input (src/parse/parse.f90, parse_source.f90, parse_detectors.f90),
read by the library's parser (`smcrt_config_load`), and the TOML-driven run: the host side of
`pathlength_scatter` (src/kernelsMod.f90:165-333) as the Fortran driver
(signedmcrt_amd/fortran/smcrt_main.f90) does it.

    cfg = load("input.toml")            # dict of the parsed tables (reference defaults applied)
    scene, cfg = scene_from_toml("input.toml")
    out = run_toml("input.toml", outdir=".")   # writes data/jmean/<fluence>, data/detectors/*
"""
from __future__ import annotations

import ctypes as C
import os
import time
from typing import Tuple

import numpy as np

from . import _abi as A
from . import detectors as D
from . import scenes
from .scene import Scene, cart_grid, source


def _s(b: bytes) -> str:
    return b.decode()


def load(path: str) -> dict:
    """Parse a TOML input file; raises SmcrtError with the reference's message on bad input."""
    cfg = A.Config()
    A.check(A.lib().smcrt_config_load(os.fsencode(path), C.byref(cfg)))
    v3 = lambda a: [float(x) for x in a]
    return {
        "source": {"name": _s(cfg.source_name), "nphotons": int(cfg.nphotons),
                   "position": v3(cfg.position) if cfg.has_position else None,
                   "direction": v3(cfg.direction) if cfg.has_direction else None,
                   "point1": v3(cfg.point1), "point2": v3(cfg.point2), "point3": v3(cfg.point3),
                   "radius": cfg.radius, "beta": cfg.beta, "radius_hi": cfg.radius_hi,
                   "annulus_type": _s(cfg.annulus_type)},
        "grid": {"nxg": cfg.nxg, "nyg": cfg.nyg, "nzg": cfg.nzg, "xmax": cfg.xmax, "ymax": cfg.ymax,
                 "zmax": cfg.zmax, "units": _s(cfg.units)},
        "geometry": {"geom_name": _s(cfg.geom_name), "tau": cfg.tau, "num_spheres": cfg.num_spheres,
                     "musb": cfg.musb, "muab": cfg.muab, "musc": cfg.musc, "muac": cfg.muac,
                     "hgg": cfg.hgg},
        "output": {"fluence": _s(cfg.fluence), "absorb": _s(cfg.absorb), "render": _s(cfg.render),
                   "render_geom": bool(cfg.render_geom), "overwrite": bool(cfg.overwrite),
                   "render_size": list(cfg.render_size)},
        "simulation": {"iseed": int(cfg.iseed), "tev": bool(cfg.tev), "absorb": bool(cfg.absorb_flag),
                       "load_checkpoint": bool(cfg.load_checkpoint),
                       "checkpoint_file": _s(cfg.checkpoint_file),
                       "checkpoint_every_n": int(cfg.checkpoint_every_n)},
        "detectors": [D.from_c(cfg.detectors[i]) for i in range(cfg.n_detectors)],
        "history_file": _s(cfg.history_file),
    }


def scene_from_toml(path: str) -> Tuple[Scene, dict]:
    """The SDF array of the registered `geom_name` (the [geometry] values its reference setup
    routine reads, scenes.TOML_PARAMS), with the file's grid, source and detectors."""
    cfg = load(path)
    geo = cfg["geometry"]
    name = geo["geom_name"]
    if name not in scenes.REGISTRY:
        raise ValueError(f"no such routine: {name!r}")
    fn = scenes.REGISTRY[name]
    params = {k: geo[k] for k in scenes.TOML_PARAMS.get(name, ())}
    arr, _, _, _ = fn(**params)
    g = cfg["grid"]
    grid = cart_grid(g["nxg"], g["nyg"], g["nzg"], g["xmax"], g["ymax"], g["zmax"])
    s = cfg["source"]
    src = source(s["name"], position=s["position"] or (0.0, 0.0, 0.0),
                 direction=s["direction"] or (0.0, 0.0, -1.0), point1=s["point1"],
                 point2=s["point2"], point3=s["point3"], radius=s["radius"])
    return Scene(arr, grid, src, cfg["detectors"]), cfg


def run_toml(path: str, outdir: str = ".", n_packets: int = 0, estimator: str = "pathlength") -> dict:
    """Run the input file end to end on the GPU and write the reference's output files under
    `outdir`: data/jmean/<fluence> (normalised FP32 NRRD) and data/detectors/detector_<i>.dat.
    `estimator`: "pathlength" (`pathlength_scatter`) or "weight" (`weight_scatter`)."""
    from .engine import Engine
    if estimator not in ("pathlength", "weight"):
        raise ValueError(f"unknown estimator {estimator!r}")
    scene, cfg = scene_from_toml(path)
    n = n_packets or cfg["source"]["nphotons"]
    eng = Engine(scene)
    t0 = time.perf_counter()
    r = eng.run(n, seed=cfg["simulation"]["iseed"], flags=A.FLAG_WEIGHT if estimator == "weight" else 0)
    secs = time.perf_counter() - t0
    L = A.lib()
    g = scene.grid
    flu = np.zeros(g.nxg * g.nyg * g.nzg, dtype=np.float32)
    raw = np.ascontiguousarray(r.jmean.reshape(-1, order="F"))
    A.check(L.smcrt_normalise_fluence(C.byref(scene.desc.grid), raw.ctypes.data_as(C.POINTER(C.c_double)),
                                      n, flu.ctypes.data_as(C.POINTER(C.c_float))))
    os.makedirs(os.path.join(outdir, "data", "jmean"), exist_ok=True)
    os.makedirs(os.path.join(outdir, "data", "detectors"), exist_ok=True)
    meta = (f"grid_data: fluence map\nnphotons: {n}\nsource: {cfg['source']['name']}\n"
            f"experiment: {cfg['geometry']['geom_name']}")
    fpath = os.path.join(outdir, "data", "jmean", cfg["output"]["fluence"])
    A.check(L.smcrt_write_nrrd(os.fsencode(fpath), flu.ctypes.data_as(C.POINTER(C.c_float)),
                               g.nxg, g.nyg, g.nzg, meta.encode()))
    dpaths = []
    for i, (d, counts) in enumerate(zip(scene.detectors, r.detectors or []), start=1):
        p = os.path.join(outdir, "data", "detectors", f"detector_{i}.dat")
        c = np.ascontiguousarray(counts.reshape(-1, order="F"), dtype=np.uint64)
        dc = d.to_c()
        A.check(L.smcrt_write_detector(os.fsencode(p), C.byref(dc), c.ctypes.data_as(C.POINTER(C.c_uint64))))
        dpaths.append(p)
    rpath = None
    if cfg["output"]["render_geom"]:  # src/kernelsMod.f90:532-536
        img = eng.render(cfg["output"]["render_size"])
        rpath = os.path.join(outdir, "data", cfg["output"]["render"])
        flat = np.ascontiguousarray(img.reshape(-1, order="F"))
        A.check(L.smcrt_write_nrrd(os.fsencode(rpath), flat.ctypes.data_as(C.POINTER(C.c_float)),
                                   *[int(v) for v in img.shape], b""))
    return {"result": r, "fluence_path": fpath, "detector_paths": dpaths, "render_path": rpath, "seconds": secs,
            "photons_per_s": n / secs, "mean_scatters": r.stats["scatters"] / n,
            "fluence": flu.reshape(g.shape, order="F")}
