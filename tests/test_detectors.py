"""Detectors (src/detectors/*.f90): the oracle's hit tests and binning pinned on hand-worked
geometry (CPU); on the GPU, the detector histograms of the same packets against the oracle
(common random numbers), and the transport unchanged by the presence of detectors."""
from __future__ import annotations

import os

import numpy as np
import pytest

import pyoracle as O
from signedmcrt_amd import config
from signedmcrt_amd.detectors import annulus_dect, camera, circle_dect

HERE = os.path.dirname(os.path.abspath(__file__))
SLAB_DECTS = os.path.join(HERE, "golden", "toml", "slab_dects.toml")


# ---- oracle, hand-worked cases ------------------------------------------------------------
def test_circle_hits_and_bins():
    # plane z = 1 facing +z (direction normalised by the reader), radius 0.5; nbins 10, maxval 1
    # -> 11 bins of width 1/9
    d = circle_dect((0, 0, 1), (0, 0, 2), 2, 0.5, nbins=10, maxval=1.0)
    up = (0, 0, 1)
    pts = [((0.1, 0.2, 0.998), up, 2),    # t = 0.002, |v| = 0.224: hit, value 0.2236 -> nint(2.01)+1 = 3
           ((0.1, 0.2, 0.99), up, 2),     # t = 0.01 > 5e-3: no
           ((0.1, 0.2, 0.998), (0, 0, -1), 2),  # denominator < 0: no
           ((0.1, 0.2, 0.998), up, 1),    # other layer: no
           ((0.6, 0.0, 0.999), up, 2),    # outside the radius: no
           ((0.0, 0.0, 1.0), up, 2)]      # t = 0, value 0: bin 1
    c = O.detector_hits(d, [p for p, _, _ in pts], [u for _, u, _ in pts], [l for _, _, l in pts])
    want = np.zeros(11, dtype=np.uint64)
    want[2] += 1
    want[0] += 1
    np.testing.assert_array_equal(c, want)
    # values beyond maxval land in the last bin
    wide = circle_dect((0, 0, 1), (0, 0, 1), 2, 10.0, nbins=10, maxval=1.0)
    c = O.detector_hits(wide, [(3.0, 0.0, 0.999)], [up], [2])
    assert c[10] == 1 and c.sum() == 1


def test_annulus_hits_and_nint_rounding():
    # centred at the origin, 0.5 <= r <= 2; nbins 4, maxval 2 -> 5 bins of width 0.5
    d = annulus_dect((0, 0, 0), (0, 0, -1), 1, 0.5, 2.0, nbins=4, maxval=2.0)
    pts = [(0.6, 0.0, 0.0),    # r 0.6, value 0.6 -> nint(1.2)+1 = 2
           (0.0, 0.0, 0.7),    # r 0.7, value 0 -> 1
           (0.3, 0.0, 0.0),    # r 0.3 < r1: no
           (1.25, 0.0, 0.0),   # value 1.25 -> 2.5 rounds half away from zero -> 3 -> bin 4
           (0.0, 1.2, 1.9)]    # r 2.25 > r2: no
    c = O.detector_hits(d, pts, [(1, 0, 0)] * len(pts), [1] * len(pts))
    np.testing.assert_array_equal(c, np.array([1, 1, 0, 1, 0], dtype=np.uint64))


def test_camera_hits_and_reference_binning():
    # plane z = -1 through p1 with edges e1 = (2,0,0), e2 = (0,2,0); unit(e2 x e1) = (0,0,-1);
    # nbins 3 -> 4x4 bins of width maxval/4 = 0.5; bins use (final z + p1.x, final y + p1.y)
    d = camera((-1, -1, -1), (1, -1, -1), (-1, 1, -1), 2, nbins=3, maxval=2.0)
    down = (0, 0, -1)
    pts = [((0.2, 0.3, 0.5), down),   # meets the plane at (0.2, 0.3): hit; x = -0.5 -> int(-1)+1 = 0 -> 4,
                                      # y = -0.7 -> int(-1.4)+1 = 0 -> 4: bin (4, 4)
           ((0.2, 0.3, 1.6), down),   # hit; x = 0.6 -> int(1.2)+1 = 2, y -> 4: bin (2, 4)
           ((1.5, 0.0, 0.0), down),   # projection 2.5 > width: no
           ((0.2, 0.3, 0.5), (0, 0, 1))]  # moving away (t < 0): no
    c = O.detector_hits(d, [p for p, _ in pts], [u for _, u in pts], [2] * len(pts))
    want = np.zeros((4, 4), dtype=np.uint64)
    want[3, 3] = 1
    want[1, 3] = 1
    np.testing.assert_array_equal(c, want)


# ---- GPU ------------------------------------------------------------------------------------
def _close_counts(a, b, what):
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    # same packets on both sides: counts differ only where an FP32 decision flips a history
    tol = 4.0 * np.sqrt(np.maximum(a, b) + 1.0) + 2.0
    bad = np.abs(a - b) > tol
    assert not bad.any(), (what, np.nonzero(bad), a[bad], b[bad])
    assert abs(a.sum() - b.sum()) <= 4.0 * np.sqrt(max(a.sum(), 1.0)) + 3.0, (what, a.sum(), b.sum())


@pytest.mark.gpu
def test_detectors_match_oracle_common_random_numbers(gpu):
    from signedmcrt_amd.engine import Engine
    scene, cfg = config.scene_from_toml(SLAB_DECTS)
    n = 100_000
    seed = cfg["simulation"]["iseed"]
    r = Engine(scene).run(n, seed=seed)
    o = O.run(scene, n, seed=seed, tally=False)
    assert len(r.detectors) == 3
    for i, (g, c) in enumerate(zip(r.detectors, o["detectors"])):
        assert g.shape == c.shape
        _close_counts(g, c, i)
    # each detector sees something: the circle the reflected packets that move on to the
    # world's top face, the annulus packets ending in the slab near the axis, the camera the
    # transmitted ones
    assert all(int(d.sum()) > 100 for d in r.detectors), [int(d.sum()) for d in r.detectors]
    # the circle counts packets whose final position lies on the world's top face: without the
    # continuation past the grid edge (z = 0.25) they would end 0.75 cm short of it and miss
    assert int(r.detectors[0].sum()) > 0.5 * r.stats["escaped_r"]


@pytest.mark.gpu
def test_detectors_do_not_change_transport(gpu):
    from signedmcrt_amd.engine import Engine
    from signedmcrt_amd.scene import Scene
    scene, cfg = config.scene_from_toml(SLAB_DECTS)
    bare = Scene(scene.array, scene.grid, scene.source)
    a = Engine(scene).run(100_000, seed=5)
    b = Engine(bare).run(100_000, seed=5)
    for k in ("absorbed", "escaped_r", "escaped_t", "killed_surface", "scatters", "ev_crossings"):
        assert a.stats[k] == b.stats[k], k
    np.testing.assert_allclose(a.jmean, b.jmean, rtol=2e-6, atol=1e-9)
    assert b.detectors is None
