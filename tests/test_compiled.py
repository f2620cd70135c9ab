"""Packet kernels compiled at run time for one scene's SDF program (hipRTC; straight-line region
code instead of the interpreter) against the library's interpreting kernels on the same packets.

The compiled code calls the same device functions on the same float constants, so packets follow
the same histories; the compiler may fuse multiply-adds differently around constants, so a rare
history can branch differently: counters agree to 1e-3 relative, the fluence sum to 1e-3."""
from __future__ import annotations

import numpy as np
import pytest

from signedmcrt_amd import scenes
from signedmcrt_amd.scene import Scene, cart_grid, source
from signedmcrt_amd.sdf import box, mono

pytestmark = pytest.mark.gpu

KEYS = ("absorbed", "escaped_r", "escaped_t", "killed_surface", "scatters", "ev_crossings")


def _close(a, b, rel=1e-3, absolute=20):
    return abs(a - b) <= max(rel * max(abs(a), abs(b)), absolute)


class _Mode:
    """Engine whose launches run with SMCRT_JIT set to `mode` (read at each launch)."""

    def __init__(self, sc, mode, monkeypatch):
        from signedmcrt_amd.engine import Engine
        self.eng, self.mode, self.mp = Engine(sc), mode, monkeypatch

    def __getattr__(self, name):
        f = getattr(self.eng, name)
        if not callable(f):
            return f

        def call(*a, **kw):
            self.mp.setenv("SMCRT_JIT", self.mode)
            return f(*a, **kw)
        return call


def _pair(sc, monkeypatch, tmp_path):
    monkeypatch.setenv("SMCRT_CACHE_DIR", str(tmp_path))
    return _Mode(sc, "0", monkeypatch), _Mode(sc, "1", monkeypatch)


def _zoo_scene():
    from test_gpu_parity import sdf_zoo
    arr = sdf_zoo() + [box((4.0, 4.0, 4.0), mono(0.0, 0.0, 0.0, 1.0), 2)]
    # a point source off every region's surface (the origin lies on several of them)
    return Scene(arr, cart_grid(32, 32, 32, 2.0, 2.0, 2.0), source("point", (0.05, 0.13, 0.21)))


def _nested_polygons_scene(square_first: bool):
    """An L-shaped polygon and a square in its notch, extruded in a slab: disjoint polygons whose
    accelerator rectangles overlap, so a point in the square lies inside both rectangles (the
    compiled kernels' gathered polygon evaluation takes the first, the second through its full
    evaluation).  Program order: the L first, or the square first."""
    from signedmcrt_amd.sdf import extrude, polygon
    ell = np.array([(-1.0, -1.0), (1.0, -1.0), (1.0, 0.0), (0.0, 0.0), (0.0, 1.0), (-1.0, 1.0)])
    sq = np.array([(0.2, 0.2), (0.8, 0.2), (0.8, 0.8), (0.2, 0.8)])
    polys = [(ell, 0.3), (sq, 1.5)]
    if square_first:
        polys = polys[::-1]
    arr = [extrude(polygon(v, mono(50.0, mua, 0.85, 1.4), k + 1), 0.25) for k, (v, mua) in enumerate(polys)]
    arr.append(box((2.4, 2.4, 0.5), mono(50.0, 0.2, 0.85, 1.4), 3))
    arr.append(box((4.4, 4.4, 2.5), mono(0.0, 0.0, 0.0, 1.0), 4))
    src = source("uniform", direction=(0.0, 0.0, -1.0), point1=(-1.2, -1.2, 0.25 - 1e-4),
                 point2=(2.4, 0.0, 0.0), point3=(0.0, 2.4, 0.0))
    return Scene(arr, cart_grid(40, 40, 8, 1.2, 1.2, 0.25), src)


def _moved_polygons_scene():
    """Two polygons in different frames (one translated, one rotated about z): the compiled
    kernels' gathered evaluation takes each lane's first containing rectangle in its own frame,
    and the disjoint-rectangle shortcut does not apply."""
    from signedmcrt_amd.sdf import extrude, invert, polygon, rotate_z, translate
    tri = np.array([(-0.6, -0.5), (0.7, -0.4), (0.1, 0.6)])
    quad = np.array([(-0.5, -0.3), (0.5, -0.3), (0.5, 0.3), (-0.5, 0.3)])
    arr = [extrude(polygon(tri, mono(50.0, 0.4, 0.85, 1.4), 1, transform=invert(translate((-0.4, 0.3, 0.0)))), 0.25),
           extrude(polygon(quad, mono(50.0, 1.2, 0.85, 1.4), 2, transform=rotate_z(0.5)), 0.25)]
    arr.append(box((2.4, 2.4, 0.5), mono(50.0, 0.2, 0.85, 1.4), 3))
    arr.append(box((4.4, 4.4, 2.5), mono(0.0, 0.0, 0.0, 1.0), 4))
    src = source("uniform", direction=(0.0, 0.0, -1.0), point1=(-1.2, -1.2, 0.25 - 1e-4),
                 point2=(2.4, 0.0, 0.0), point3=(0.0, 2.4, 0.0))
    return Scene(arr, cart_grid(40, 40, 8, 1.2, 1.2, 0.25), src)


@pytest.mark.parametrize("name", ["refracting_sphere", "skin_vessels", "polygon_regions", "zoo",
                                  "nested_polygons", "nested_polygons_square_first", "moved_polygons"])
def test_compiled_kernels_match_interpreter(gpu, name, monkeypatch, tmp_path):
    if name == "zoo":
        sc = _zoo_scene()
    elif name.startswith("nested_polygons"):
        sc = _nested_polygons_scene(name.endswith("square_first"))
    elif name == "moved_polygons":
        sc = _moved_polygons_scene()
    else:
        kw = {"nxyz": 32} if name == "skin_vessels" else ({"n_side": 2, "n_edges": 40, "nxy": 40, "nz": 8}
                                                          if name == "polygon_regions" else {})
        sc, _ = scenes.build(name, **kw)
    interp, comp = _pair(sc, monkeypatch, tmp_path)
    a = interp.run(300_000, seed=23)
    b = comp.run(300_000, seed=23)
    assert "compiled" in comp.kernel_variant()[0], comp.kernel_variant()
    assert "compiled" not in interp.kernel_variant()[0]
    for k in KEYS:
        assert _close(a.stats[k], b.stats[k]), (k, a.stats[k], b.stats[k])
    assert b.stats["err_unreachable"] == 0 and b.stats["err_nan"] == 0
    # (the overlapping zoo regions kill most packets at interfaces, on the oracle too)
    assert b.stats["killed_surface"] < b.stats["launched"] and b.stats["scatters"] > 0
    assert abs(a.jmean.sum() / b.jmean.sum() - 1) < 1e-3
    m = a.jmean > 0.05 * a.jmean.max()
    assert np.median(np.abs(a.jmean[m] / b.jmean[m] - 1)) < 1e-3
    # 2-, 3- and 4-wave compiled builds: the same packets
    for w in (2, 4):
        c = comp.run(300_000, seed=23, waves=w)
        for k in KEYS:
            assert c.stats[k] == b.stats[k], (w, k)


def test_compiled_adjoint_matches_interpreter(gpu, monkeypatch, tmp_path):
    sc, _ = scenes.build("skin_vessels", nxyz=32)
    interp, comp = _pair(sc, monkeypatch, tmp_path)
    ga = interp.grad(100_000, seed=29).d_w_dmua
    gb = comp.grad(100_000, seed=29).d_w_dmua
    assert "compiled" in comp.kernel_variant()[0]
    np.testing.assert_allclose(gb, ga, rtol=2e-3, atol=1e-3 * np.abs(ga).max())


def test_compiled_kernels_are_cached(gpu, monkeypatch, tmp_path):
    """A second scene handle with the same program loads the code object from the disk cache."""
    from signedmcrt_amd.engine import Engine
    monkeypatch.setenv("SMCRT_CACHE_DIR", str(tmp_path))
    monkeypatch.setenv("SMCRT_JIT", "1")
    sc, _ = scenes.build("refracting_sphere", nxyz=32)
    e1 = Engine(sc)
    r1 = e1.run(50_000, seed=3, tally=False)
    assert e1.kernel_variant()[0] == "basic+compiled"
    assert len(list((tmp_path / "jit").glob("*.co"))) == 1
    e2 = Engine(sc)
    r2 = e2.run(50_000, seed=3, tally=False)
    assert e2.kernel_variant()[0] == "basic+compiled(cached)"
    for k in KEYS:
        assert r1.stats[k] == r2.stats[k], k


def test_compiled_kernels_with_detectors(gpu, monkeypatch, tmp_path):
    """A scene with detectors compiles its kernels with the detector records and the ghost
    continuation; detector counts agree with the interpreter's."""
    import os
    from signedmcrt_amd import config
    toml = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "toml", "slab_dects.toml")
    sc, _ = config.scene_from_toml(toml)
    interp, comp = _pair(sc, monkeypatch, tmp_path)
    a = interp.run(400_000, seed=7)
    b = comp.run(400_000, seed=7)
    assert "compiled" in comp.kernel_variant()[0]
    for k in KEYS:
        assert _close(a.stats[k], b.stats[k]), (k, a.stats[k], b.stats[k])
    for da, db in zip(a.detectors, b.detectors):
        assert da.sum() > 0
        assert _close(float(da.sum()), float(db.sum())), (da.sum(), db.sum())


def test_diagnostic_launches_use_the_library_kernels(gpu, monkeypatch, tmp_path):
    """Event traces and scatter moments are left out of the compiled kernels: launches asking for
    them run the library's kernels even with SMCRT_JIT=1."""
    monkeypatch.setenv("SMCRT_CACHE_DIR", str(tmp_path))
    monkeypatch.setenv("SMCRT_JIT", "1")
    from signedmcrt_amd.engine import Engine
    sc, _ = scenes.build("scat_test2")
    eng = Engine(sc)
    r = eng.run(1000, seed=5, tally=False, trace_packets=10, trace_events=4)
    assert "compiled" not in eng.kernel_variant()[0]
    assert np.all(r.trace[:, 0, 3] == 1)
