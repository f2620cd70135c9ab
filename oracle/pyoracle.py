"""ctypes binding of the CPU oracle (oracle/liboracle.so) and the flang-built reference shim.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsmcrt_ref.so")

_P = C.POINTER
_dp = _P(C.c_double)


def _abi():
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from signedmcrt_amd import _abi as A  # ctypes struct layouts only (no library load)
    return A


A = _abi()


class OrcOpts(C.Structure):
    _fields_ = [("n_packets", C.c_uint64), ("seed", C.c_uint64), ("first_packet", C.c_uint64),
                ("flags", C.c_uint32), ("max_bounces", C.c_int32), ("threads", C.c_int32),
                ("trace_packets", C.c_int32), ("trace_events", C.c_int32), ("grad", C.c_int32),
                ("weights", _P(C.c_float))]


class OrcOut(C.Structure):
    _fields_ = [("jmean", _dp), ("trace", _P(C.c_float)),
                ("d_r_dmua", _dp), ("d_r_dmus", _dp), ("d_w_dmua", _dp), ("d_w_dmus", _dp),
                ("w_tally", C.c_double), ("stats", A.Stats), ("seconds", C.c_double),
                ("det_counts", _P(C.c_uint64)), ("weight_deposit", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(ORACLE_SO)
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [_P(A.SceneDesc), _P(OrcOpts), _P(OrcOut)]
        L.orc_eval.restype = None
        L.orc_eval.argtypes = [_P(A.SceneDesc), C.c_int, C.c_int64, _dp, _dp, _dp]
        L.orc_fresnel.restype = None
        L.orc_fresnel.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_scatter.restype = None
        L.orc_scatter.argtypes = [C.c_int64, _dp, C.c_double, C.c_uint64, C.c_uint64, _dp]
        L.orc_emit.restype = None
        L.orc_emit.argtypes = [_P(A.SceneDesc), C.c_int64, C.c_uint64, C.c_uint64, _dp, _dp]
        L.orc_walk.restype = C.c_double
        L.orc_walk.argtypes = [_P(A.SceneDesc), _dp, _dp, C.c_double, _dp, _P(C.c_int)]
        L.orc_detector_hits.restype = None
        L.orc_detector_hits.argtypes = [_P(A.Detector), C.c_int64, _dp, _dp, _P(C.c_int), _P(C.c_uint64)]
        L.orc_render.restype = None
        L.orc_render.argtypes = [_P(A.SceneDesc), _P(C.c_int32), _dp]
        L.orc_uniforms.restype = None
        L.orc_uniforms.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, _dp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct=C.c_double):
    return a.ctypes.data_as(_P(ct))


def evaluate(scene, region: int, pts, normals: bool = False):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    d = np.empty(n)
    nn = np.empty((n, 3)) if normals else None
    lib().orc_eval(C.byref(scene.desc), region, n, _ptr(pts), _ptr(d),
                   _ptr(nn) if normals else None)
    return (d, nn) if normals else d


def fresnel(I, N, n1, n2):
    I = np.ascontiguousarray(I, dtype=np.float64).reshape(-1, 3)
    N = np.ascontiguousarray(N, dtype=np.float64).reshape(-1, 3)
    n = len(I)
    n1 = np.ascontiguousarray(np.broadcast_to(n1, (n,)), dtype=np.float64)
    n2 = np.ascontiguousarray(np.broadcast_to(n2, (n,)), dtype=np.float64)
    R = np.empty(n)
    refl = np.empty((n, 3))
    refr = np.empty((n, 3))
    lib().orc_fresnel(n, _ptr(I), _ptr(N), _ptr(n1), _ptr(n2), _ptr(R), _ptr(refl), _ptr(refr))
    return R, refl, refr


def scatter(dirs, g: float, seed: int = 1, id0: int = 0):
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(dirs)
    lib().orc_scatter(len(dirs), _ptr(dirs), float(g), seed, id0, _ptr(out))
    return out


def emit(scene, n: int, seed: int = 1, id0: int = 0):
    pos = np.empty((n, 3))
    dirs = np.empty((n, 3))
    lib().orc_emit(C.byref(scene.desc), n, seed, id0, _ptr(pos), _ptr(dirs))
    return pos, dirs


def walk(scene, pos, dir, dstep: float):
    g = scene.grid
    jm = np.zeros(g.nxg * g.nyg * g.nzg)
    p = np.ascontiguousarray(pos, dtype=np.float64)
    d = np.ascontiguousarray(dir, dtype=np.float64)
    flag = C.c_int(0)
    t = lib().orc_walk(C.byref(scene.desc), _ptr(p), _ptr(d), dstep, _ptr(jm), C.byref(flag))
    return t, jm.reshape(g.shape, order="F"), bool(flag.value)


def detector_hits(det, pos, dirs, layers):
    """Bins of detector `det` (signedmcrt_amd.detectors description) after the given hit points."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    lay = np.ascontiguousarray(np.broadcast_to(layers, (len(pos),)), dtype=np.int32)
    counts = np.zeros(int(np.prod(det.shape)), dtype=np.uint64)
    dc = det.to_c()
    lib().orc_detector_hits(C.byref(dc), len(pos), _ptr(pos), _ptr(dirs), _ptr(lay, C.c_int),
                            _ptr(counts, C.c_uint64))
    return counts.reshape(det.shape, order="F")


def render(scene, samples):
    """`render_sub` in FP64: the scene's minimum signed distance on the render grid."""
    s = (C.c_int32 * 3)(*[int(v) for v in samples])
    img = np.empty(int(np.prod(samples)))
    lib().orc_render(C.byref(scene.desc), s, _ptr(img))
    return img.reshape(tuple(int(v) for v in samples), order="F")


def uniforms(seed: int, pid: int, n: int):
    out = np.empty(n)
    lib().orc_uniforms(seed, pid, n, _ptr(out))
    return out


def run(scene, n_packets: int, seed: int = 123456789, first_packet: int = 0, flags: int = 0,
        threads: int = 0, tally: bool = True, trace_packets: int = 0, trace_events: int = 0,
        grad: bool = False, weights=None, max_bounces: int = 0):
    """Run the oracle's packet loop.  Returns a dict: jmean (raw FP64, (nx,ny,nz) Fortran
    order) or None, stats, seconds, and the pMC gradients when grad=True."""
    g = scene.grid
    o = OrcOpts()
    o.n_packets, o.seed, o.first_packet, o.flags = n_packets, seed, first_packet, flags
    o.max_bounces, o.threads = max_bounces, threads
    o.trace_packets, o.trace_events, o.grad = trace_packets, trace_events, int(grad)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape(-1, order="F"))
        o.weights = _ptr(w, C.c_float)
    out = OrcOut()
    jm = np.zeros(g.nxg * g.nyg * g.nzg) if tally else None
    if tally:
        out.jmean = _ptr(jm)
    tr = None
    if trace_packets:
        tr = np.zeros(trace_packets * trace_events * 4, dtype=np.float32)
        out.trace = _ptr(tr, C.c_float)
    nm = scene.n_materials
    gr = [np.zeros(nm) for _ in range(4)]
    if grad:
        out.d_r_dmua, out.d_r_dmus, out.d_w_dmua, out.d_w_dmus = (_ptr(x) for x in gr)
    dets = getattr(scene, "detectors", [])
    sizes = [int(np.prod(dd.shape)) for dd in dets]
    counts = np.zeros(max(1, sum(sizes)), dtype=np.uint64)
    if dets:
        out.det_counts = _ptr(counts, C.c_uint64)
    rc = lib().orc_run(C.byref(scene.desc), C.byref(o), C.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_run failed ({rc})")
    res = {"jmean": jm.reshape(g.shape, order="F") if tally else None,
           "stats": out.stats.as_dict(), "seconds": out.seconds,
           "weight_deposit": out.weight_deposit}
    if dets:
        res["detectors"], off = [], 0
        for dd, sz in zip(dets, sizes):
            res["detectors"].append(counts[off:off + sz].reshape(dd.shape, order="F").copy())
            off += sz
    if trace_packets:
        res["trace"] = tr.reshape(trace_packets, trace_events, 4)
    if grad:
        res.update(d_r_dmua=gr[0], d_r_dmus=gr[1], d_w_dmua=gr[2], d_w_dmus=gr[3],
                   w_tally=out.w_tally)
    return res


# ---- the flang-built reference routines (oracle/_ref) ----------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.ref_reflect_refract.restype = None
        L.ref_reflect_refract.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_int, _dp,
                                          _P(C.c_int), _dp]
        L.ref_transform.restype = None
        L.ref_transform.argtypes = [C.c_int64, _dp, _dp, _dp]
        L.ref_grid_faces.restype = None
        L.ref_grid_faces.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, _dp, _dp, _dp]
        _ref = L
    return _ref


def ref_reflect_refract(I, N, n1, n2, seed: int = 123456789):
    I = np.ascontiguousarray(I, dtype=np.float64).reshape(-1, 3)
    N = np.ascontiguousarray(N, dtype=np.float64).reshape(-1, 3)
    n = len(I)
    n1 = np.ascontiguousarray(np.broadcast_to(n1, (n,)), dtype=np.float64)
    n2 = np.ascontiguousarray(np.broadcast_to(n2, (n,)), dtype=np.float64)
    R = np.empty(n)
    fl = np.empty(n, dtype=np.int32)
    out = np.empty((n, 3))
    ref_lib().ref_reflect_refract(n, _ptr(I), _ptr(N), _ptr(n1), _ptr(n2), seed, _ptr(R),
                                  fl.ctypes.data_as(_P(C.c_int)), _ptr(out))
    return R, fl.astype(bool), out


def ref_transform(T, pts):
    T = np.asfortranarray(np.asarray(T, dtype=np.float64))
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(pts)
    ref_lib().ref_transform(len(pts), _ptr(T), _ptr(pts), _ptr(out))
    return out


def ref_grid_faces(nx, ny, nz, xmax, ymax, zmax):
    xf, yf, zf = np.empty(nx + 1), np.empty(ny + 1), np.empty(nz + 2)
    ref_lib().ref_grid_faces(nx, ny, nz, xmax, ymax, zmax, _ptr(xf), _ptr(yf), _ptr(zf))
    return xf, yf, zf
