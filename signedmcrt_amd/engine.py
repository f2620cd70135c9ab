"""Run driver over libsmcrt.so: the host side of `pathlength_scatter` (src/kernelsMod.f90:165-333).

`Engine(scene)` owns one `smcrt_scene` handle.  `run()` is the packet loop (the reference's
OpenMP region) returning the raw FP64 path-length grid and the termination/event counters;
`evaluate()` is `sdf%evaluate`/`calcNormal`; `grad()` is the perturbation-MC adjoint.
Everything here calls the HIP library; nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _abi as A
from .scene import Scene


@dataclass
class RunResult:
    jmean: Optional[np.ndarray]  # raw FP64 path-length sums, shape (nx, ny, nz), Fortran order
    stats: dict
    trace: Optional[np.ndarray] = None
    detectors: Optional[list] = None  # per detector: its bins (see signedmcrt_amd.detectors)

    @property
    def n(self) -> int:
        return self.stats["launched"]

    def fractions(self) -> dict:
        n = max(self.n, 1)
        s = self.stats
        return {"A": s["absorbed"] / n, "R": s["escaped_r"] / n, "T": s["escaped_t"] / n,
                "K": (s["killed_surface"] + s["killed_bounce"]) / n}


@dataclass
class GradResult:
    d_r_dmua: np.ndarray
    d_r_dmus: np.ndarray
    d_w_dmua: np.ndarray
    d_w_dmus: np.ndarray
    r_d: float
    w_tally: float
    stats: dict
    jmean: Optional[np.ndarray] = None


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def run_opts(n: int, seed: int = 123456789, first_packet: int = 0, flags: int = 0,
             block: int = 0, blocks_per_cu: int = 0, chunk: int = 0, max_bounces: int = 0,
             trace_packets: int = 0, trace_events: int = 0, walk_cap: int = 0,
             refill_min: int = 0, waves: int = 0) -> A.RunOpts:
    o = A.RunOpts()
    o.n_packets, o.seed, o.first_packet, o.flags = int(n), int(seed), int(first_packet), int(flags)
    o.block_size, o.blocks_per_cu, o.chunk, o.max_bounces = block, blocks_per_cu, chunk, max_bounces
    o.trace_packets, o.trace_events, o.walk_cap = trace_packets, trace_events, walk_cap
    o.refill_min, o.waves = refill_min, waves
    return o


class Engine:
    def __init__(self, scene: Scene, shape: Optional[dict] = None):
        """`shape`: launch shape applied to every run (an autotune result, see `tune`)."""
        self.scene = scene
        self.shape = dict(shape or {})
        self._lib = A.lib()
        h = C.c_void_p()
        A.check(self._lib.smcrt_scene_create(C.byref(scene.desc), C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self._lib.smcrt_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, flags: int, shape: dict):
        """Explicit keywords win over the engine's tuned shape; flags combine (an explicit partial-grid
        count replaces the tuned one)."""
        if not self.shape:
            return flags, shape
        from .autotune import run_kwargs
        kw = run_kwargs(self.shape)
        tuned = kw.pop("flags")
        if flags & A.FLAG_GRID_COPIES_MASK:
            tuned &= ~A.FLAG_GRID_COPIES_MASK
        flags |= tuned
        kw.update(shape)
        return flags, kw

    def tune(self, **kw):
        """Autotune the launch shape on the current GPU (cached per scene and device) and use it
        for later runs.  Returns the autotune.TuneResult."""
        from . import autotune
        res = autotune.tune(self, **kw)
        self.shape = dict(res.shape)
        return res

    # -- the packet loop -------------------------------------------------------------------
    def run(self, n: int, seed: int = 123456789, first_packet: int = 0, flags: int = 0,
            tally: bool = True, trace_packets: int = 0, trace_events: int = 0,
            **shape) -> RunResult:
        g = self.scene.grid
        flags, shape = self._shape(flags, shape)
        o = run_opts(n, seed, first_packet, flags, trace_packets=trace_packets,
                     trace_events=trace_events, **shape)
        t = A.Tallies()
        jm = None
        if tally:
            jm = np.zeros(g.nxg * g.nyg * g.nzg, dtype=np.float64)
            t.jmean = _dptr(jm)
        tr = None
        if trace_packets:
            tr = np.zeros(trace_packets * trace_events * 4, dtype=np.float32)
            t.trace = _fptr(tr)
        dets = self.scene.detectors
        counts = None
        if dets:
            nbins = [int(self._lib.smcrt_detector_bins(self.handle, i)) for i in range(len(dets))]
            counts = np.zeros(sum(nbins), dtype=np.uint64)
            t.detector_counts = counts.ctypes.data_as(C.POINTER(C.c_uint64))
        A.check(self._lib.smcrt_run(self.handle, C.byref(o), C.byref(t)))
        det_out = None
        if dets:
            det_out, off = [], 0
            for d, nb in zip(dets, nbins):
                det_out.append(counts[off:off + nb].reshape(d.shape, order="F").copy())
                off += nb
        return RunResult(jm.reshape(g.shape, order="F") if tally else None, t.stats.as_dict(),
                         tr.reshape(trace_packets, trace_events, 4) if trace_packets else None,
                         det_out)

    def run_device(self, n: int, d_jmean_ptr: int, seed: int = 123456789, first_packet: int = 0,
                   flags: int = 0, stream: int = 0, **shape) -> dict:
        """Packet loop into a device-resident FP64 grid (a raw device pointer, e.g.
        torch.Tensor.data_ptr()); returns the counters."""
        flags, shape = self._shape(flags, shape)
        o = run_opts(n, seed, first_packet, flags, **shape)
        st = A.Stats()
        A.check(self._lib.smcrt_run_device(self.handle, C.byref(o), C.c_void_p(d_jmean_ptr),
                                           C.byref(st), C.c_void_p(stream)))
        return st.as_dict()

    def kernel_variant(self):
        """(name, scene kind mask, variant kind mask) of the packet kernel this scene runs."""
        sm, vm = C.c_uint32(), C.c_uint32()
        name = self._lib.smcrt_scene_kernel(self.handle, C.byref(sm), C.byref(vm))
        return name.decode(), sm.value, vm.value

    # -- SDF evaluation ----------------------------------------------------------------------
    def evaluate(self, region: int, pts, normals: bool = False):
        p = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        n = len(p)
        d = np.empty(n, dtype=np.float32)
        nn = np.empty((n, 3), dtype=np.float32) if normals else None
        A.check(self._lib.smcrt_evaluate(self.handle, region, n, _fptr(p), _fptr(d),
                                         _fptr(nn) if normals else None))
        return (d, nn) if normals else d

    def render(self, samples=(200, 200, 200)) -> np.ndarray:
        """`render_sub` (src/sdfs/sdf_base.f90:302-346): the scene's minimum signed distance on a
        samples[0] x samples[1] x samples[2] grid (FP32, Fortran order)."""
        s = (C.c_int32 * 3)(*[int(v) for v in samples])
        img = np.empty(int(np.prod(samples)), dtype=np.float32)
        A.check(self._lib.smcrt_render(self.handle, s, _fptr(img)))
        return img.reshape(tuple(int(v) for v in samples), order="F")

    # -- materials / adjoint -------------------------------------------------------------------
    def set_materials(self, mus=None, mua=None) -> None:
        arr = self.scene.material_array(mus, mua)
        A.check(self._lib.smcrt_scene_set_materials(self.handle, arr, len(self.scene.materials)))

    def grad(self, n: int, seed: int = 123456789, first_packet: int = 0, weights=None,
             tally: bool = False, materials=None, **shape) -> GradResult:
        """pMC adjoint.  `materials`: indices of the materials to differentiate (default all;
        at most 16 — the others' derivatives come back as 0)."""
        g = self.scene.grid
        nm = self.scene.n_materials
        # the adjoint kernel's own tuned resident waves, unless given
        if "waves" not in shape and self.shape.get("adj_waves"):
            shape = dict(shape, waves=int(self.shape["adj_waves"]))
        flags, shape = self._shape(0, shape)
        o = run_opts(n, seed, first_packet, flags, **shape)
        if materials is not None:
            mask = 0
            for m in materials:
                if not 0 <= int(m) < nm:
                    raise ValueError(f"material {m} out of range")
                mask |= 1 << int(m)
            o.grad_mask_lo, o.grad_mask_hi = mask & 0xFFFFFFFF, mask >> 32
        out = A.GradOut()
        arrs = [np.zeros(nm) for _ in range(4)]
        out.d_r_dmua, out.d_r_dmus, out.d_w_dmua, out.d_w_dmus = (_dptr(a) for a in arrs)
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape(-1, order="F"))
            if w.size != g.nxg * g.nyg * g.nzg:
                raise ValueError("weights must have one value per voxel")
        jm = np.zeros(g.nxg * g.nyg * g.nzg) if tally else None
        A.check(self._lib.smcrt_grad(self.handle, C.byref(o), _fptr(w) if w is not None else None,
                                     _dptr(jm) if tally else None, C.byref(out)))
        return GradResult(arrs[0], arrs[1], arrs[2], arrs[3], float(out.r_d), float(out.w_tally),
                          out.stats.as_dict(), jm.reshape(g.shape, order="F") if tally else None)
