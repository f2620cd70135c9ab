"""Independent pins of the oracle functions the reference's own test vectors leave open — CPU only.

Each check is a closed form, an identity between two SDFs, a textbook formula or a distribution
property that a plausible slip (a dropped term, a wrong sign, a swapped axis, h^2 for h^3) fails:
  - SmoothUnion (src/sdfs/sdfModifiers.f90:427-442): equals min(d1, d2) once |d1 - d2| >= k,
    min - k/6 where d1 = d2, lies in [min - k/6, min], and its value, gradient and curvature
    are continuous across |d1 - d2| = k (the cubic blend; a quadratic one has a curvature jump);
  - onion (:307-316): ||p| - R| - t for a sphere;
  - twist (:335-355): k = 0 is the identity, a z-axis cylinder is invariant, and a twisted box is
    the box evaluated at the point rotated by k z about z;
  - elongate (:318-333): a sphere elongated along x is the capsule of that segment, and along all
    three axes the box of those half-sizes minus the radius;
  - revolution (:290-305): a revolved sphere is the torus about the y axis, a revolved box the
    distance to a rectangle in the (sqrt(x^2 + z^2), y) half-plane;
  - the circular source (src/photon.f90:214-287): the reference's own property
    (test/photon/test_photon.f90:195-210), the disc perpendicular to the direction and centred at
    -position, uniform over the disc (E r^2 = R^2/2, r^2/R^2 and the angle uniform);
  - the uniform source (src/photon.f90:397-461): pos = p1 + u p2 + v p3 with u, v uniform on (0,1);
  - the interface table and Fresnel events (src/inttau2.f90:104-185, src/surfaces.f90): a clear
    n = 1.5 slab in air under a collimated beam gives killed = R + (1 - R) R and transmitted
    (1 - R)^2 with R from the textbook Fresnel equations, the beam leaves parallel to the way it
    came and refracts by Snell's law inside; from inside beyond the critical angle every packet is
    reflected (and killed); a crossing at 60 degrees always ends within eps of the surface and is
    killed by the next flight start.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import stats

import pyoracle as O
from pin_scenes import clear_slab, clear_slab_expected, fresnel_textbook, ray_directions
from signedmcrt_amd.scene import Scene, cart_grid, source
from signedmcrt_amd.sdf import (SmoothUnion, box, capsule, cylinder, elongate, invert, model, mono,
                                onion, revolution, sphere, translate, twist, union)

M0 = mono(0.0, 0.0, 0.0, 0.0)


def _scene(arr, half=5.0, src=None):
    return Scene(arr, cart_grid(4, 4, 4, half, half, half), src or source("point"))


def _ev(sdf, pts, half=5.0):
    return O.evaluate(_scene([sdf], half), 0, np.asarray(pts, dtype=float))


# ---- SmoothUnion ------------------------------------------------------------------------------
K = 0.3


def _two_spheres(op, k=K):
    a = sphere(0.5, M0, 1)
    b = sphere(0.5, M0, 1, transform=invert(translate((1.2, 0.0, 0.0))))
    return model([a, b], op, k)


def _d12(p):
    d1 = np.linalg.norm(p, axis=1) - 0.5
    d2 = np.linalg.norm(p - [1.2, 0.0, 0.0], axis=1) - 0.5
    return d1, d2


def test_smooth_union_limits():
    rng = np.random.default_rng(11)
    p = rng.uniform(-1.5, 2.5, (20000, 3))
    d1, d2 = _d12(p)
    got = _ev(_two_spheres(SmoothUnion), p)
    far = np.abs(d1 - d2) >= K
    # outside the blend band it is the plain minimum
    np.testing.assert_allclose(got[far], np.minimum(d1, d2)[far], atol=1e-13)
    # everywhere between min - k/6 and min, strictly below min inside the band
    assert np.all(got <= np.minimum(d1, d2) + 1e-13)
    assert np.all(got >= np.minimum(d1, d2) - K / 6 - 1e-13)
    assert np.all(got[~far] < np.minimum(d1, d2)[~far])
    # on the bisecting plane x = 0.6 the two distances are equal: min - k/6
    q = np.c_[np.full(200, 0.6), rng.uniform(-1, 1, (200, 2))]
    e1, _ = _d12(q)
    np.testing.assert_allclose(_ev(_two_spheres(SmoothUnion), q), e1 - K / 6, atol=1e-13)
    # union of the same pair is the minimum everywhere
    np.testing.assert_allclose(_ev(_two_spheres(union), p), np.minimum(d1, d2), atol=1e-13)


def _band_edges(y=0.3):
    """x where d1 - d2 = -K and +K on the line (x, y, 0) (bisection on the closed-form distances)."""
    out = []
    for target in (-K, K):
        lo, hi = -0.5, 1.7
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            d1, d2 = _d12(np.array([[mid, y, 0.0]]))
            lo, hi = (mid, hi) if d1[0] - d2[0] < target else (lo, mid)
        out.append(0.5 * (lo + hi))
    return out


def test_smooth_union_continuous_at_band_edge():
    """Value and gradient are continuous where |d1 - d2| crosses k (the cubic blend meets the
    minimum with matching value and slope; a k/4 or a linear blend would leave a step)."""
    for xe in _band_edges():
        h = 1e-6
        xs = xe + np.array([-2 * h, -h, h, 2 * h])
        p = np.c_[xs, np.full(4, 0.3), np.zeros(4)]
        v = _ev(_two_spheres(SmoothUnion), p)
        # one-sided slopes on each side agree, and the value has no jump
        s_lo, s_hi = (v[1] - v[0]) / h, (v[3] - v[2]) / h
        assert abs(s_lo - s_hi) < 1e-4, (xe, s_lo, s_hi)
        assert abs((v[2] - v[1]) - 0.5 * (s_lo + s_hi) * 2 * h) < 1e-9


def test_smooth_union_curvature_continuous_at_band_edge():
    """The cubic blend min - h^3 k/6 is twice continuously differentiable where it meets the
    minimum ((h^3)'' = 6 h h'^2 -> 0), so the curvature along a line has no jump at the band edge;
    a quadratic blend (h^2)'' = 2 h'^2 jumps there by about 4 on this line."""
    step = 1e-4
    for xe in _band_edges():
        curv = []
        for side in (-1, 1):
            xc = xe + side * 3 * step
            xs = xc + np.array([-step, 0.0, step])
            p = np.c_[xs, np.full(3, 0.3), np.zeros(3)]
            v = _ev(_two_spheres(SmoothUnion), p)
            curv.append((v[0] - 2 * v[1] + v[2]) / step ** 2)
        assert abs(curv[0] - curv[1]) < 0.1, (xe, curv)


# ---- onion, twist, elongate, revolution --------------------------------------------------------
def test_onion_of_sphere():
    rng = np.random.default_rng(12)
    p = rng.uniform(-1.5, 1.5, (5000, 3))
    want = np.abs(np.linalg.norm(p, axis=1) - 0.8) - 0.05
    np.testing.assert_allclose(_ev(onion(sphere(0.8, M0, 1), 0.05), p), want, atol=1e-13)


def test_twist_identities():
    rng = np.random.default_rng(13)
    p = rng.uniform(-1.5, 1.5, (5000, 3))
    b = lambda: box((1.2, 0.5, 2.0), M0, 1)
    # k = 0: the identity
    np.testing.assert_allclose(_ev(twist(b(), 0.0), p), _ev(b(), p), atol=1e-14)
    # a cylinder along z is symmetric about z: any twist leaves it unchanged
    cyl = lambda: cylinder((0.0, 0.0, -0.8), (0.0, 0.0, 0.8), 0.5, M0, 1)
    np.testing.assert_allclose(_ev(twist(cyl(), 1.7), p), _ev(cyl(), p), atol=1e-12)
    # a twisted box is the box at the point rotated by angle k z about z (x' = c x - s y,
    # y' = s x + c y); the box has no symmetry that hides a sign slip in s
    k = 1.3
    c, s = np.cos(k * p[:, 2]), np.sin(k * p[:, 2])
    q = np.c_[c * p[:, 0] - s * p[:, 1], s * p[:, 0] + c * p[:, 1], p[:, 2]]
    np.testing.assert_allclose(_ev(twist(b(), k), p), _ev(b(), q), atol=1e-13)
    qwrong = np.c_[c * p[:, 0] + s * p[:, 1], -s * p[:, 0] + c * p[:, 1], p[:, 2]]
    assert np.max(np.abs(_ev(twist(b(), k), p) - _ev(b(), qwrong))) > 0.05


def test_elongate_closed_forms():
    rng = np.random.default_rng(14)
    p = rng.uniform(-1.5, 1.5, (5000, 3))
    # a sphere elongated along x by h is the capsule from (-h,0,0) to (h,0,0) of the same radius
    got = _ev(elongate(sphere(0.3, M0, 1), (0.6, 0.0, 0.0)), p)
    np.testing.assert_allclose(got, _ev(capsule((-0.6, 0, 0), (0.6, 0, 0), 0.3, M0, 1), p), atol=1e-13)
    # along all three axes: the box of half-sizes h, rounded by r (box distance - r)
    h = np.array([0.4, 0.2, 0.3])
    got = _ev(elongate(sphere(0.25, M0, 1), h), p)
    np.testing.assert_allclose(got, _ev(box(2 * h, M0, 1), p) - 0.25, atol=1e-13)


def test_revolution_closed_forms():
    rng = np.random.default_rng(15)
    p = rng.uniform(-1.5, 1.5, (5000, 3))
    rho = np.hypot(p[:, 0], p[:, 2])
    # a revolved sphere: torus of major radius o about the y axis
    got = _ev(revolution(sphere(0.2, M0, 1), 0.7), p)
    np.testing.assert_allclose(got, np.hypot(rho - 0.7, p[:, 1]) - 0.2, atol=1e-13)
    # a revolved box (half-sizes a, b in its x, y; z half-size large enough never to bind): the
    # signed distance to the rectangle |rho - o| <= a, |y| <= b
    a, b, o = 0.15, 0.35, 0.8
    got = _ev(revolution(box((2 * a, 2 * b, 20.0), M0, 1), o), p, half=20.0)
    qx, qy = np.abs(rho - o) - a, np.abs(p[:, 1]) - b
    want = np.hypot(np.maximum(qx, 0), np.maximum(qy, 0)) + np.minimum(np.maximum(qx, qy), 0)
    np.testing.assert_allclose(got, want, atol=1e-13)


# ---- emitters -----------------------------------------------------------------------------------
def _emit(src, n=40000, seed=3):
    sc = Scene([box((20, 20, 20), M0, 1)], cart_grid(4, 4, 4, 10, 10, 10), src)
    return O.emit(sc, n, seed=seed)


@pytest.mark.parametrize("direction", [(0.0, 0.0, -1.0), (1.0, 2.0, -2.0), (1.0, 0.0, 0.0)])
def test_circular_source_disc(direction):
    """test/photon/test_photon.f90:195-210 (position (1,1,1), radius 2.5: every point within 2.5
    of (-1,-1,-1)), plus the disc's plane and uniformity, for a direction along -z, a general one
    and +x (where the reference swaps its reference axis, src/photon.f90:252-259)."""
    pos, R = np.array([1.0, 1.0, 1.0]), 2.5
    d = np.array(direction) / np.linalg.norm(direction)
    p, dirs = _emit(source("circular", position=tuple(pos), direction=tuple(d), radius=R))
    v = p + pos  # relative to the disc centre -position
    r2 = np.sum(v * v, axis=1)
    n = len(p)
    assert np.all(np.sqrt(r2) <= R * (1 + 1e-12))
    np.testing.assert_allclose(v @ d, 0.0, atol=1e-12)
    np.testing.assert_allclose(dirs, np.tile(d, (n, 1)), atol=1e-15)
    # uniform over the disc: r^2 / R^2 ~ U(0,1), E r^2 = R^2 / 2, no preferred direction
    assert stats.kstest(r2 / R ** 2, "uniform").pvalue > 1e-3
    assert abs(r2.mean() - R * R / 2) < 4 * (R * R / math.sqrt(12)) / math.sqrt(n)
    assert np.linalg.norm(v.mean(axis=0)) < 4 * R / math.sqrt(2 * n)
    # the azimuth in the disc's own frame is uniform
    e1 = np.cross(d, [0.3, 0.5, 0.7])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    ang = np.arctan2(v @ e2, v @ e1)
    assert stats.kstest((ang + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3


def test_uniform_source_parallelogram():
    p1, p2, p3 = np.array([-1.0, -0.5, 0.3]), np.array([2.0, 0.5, 0.0]), np.array([0.0, 1.5, -0.4])
    d = np.array([0.0, 0.6, -0.8])
    p, dirs = _emit(source("uniform", direction=tuple(d), point1=tuple(p1), point2=tuple(p2), point3=tuple(p3)))
    uv, *_ = np.linalg.lstsq(np.c_[p2, p3], (p - p1).T, rcond=None)
    u, v = uv
    np.testing.assert_allclose(np.c_[p2, p3] @ uv + p1[:, None], p.T, atol=1e-12)  # in the plane
    assert np.all((u > 0) & (u < 1) & (v > 0) & (v < 1))
    for w in (u, v):
        assert stats.kstest(w, "uniform").pvalue > 1e-3
    assert abs(np.corrcoef(u, v)[0, 1]) < 4 / math.sqrt(len(u))
    np.testing.assert_allclose(dirs, np.tile(d, (len(p), 1)), atol=1e-15)


# ---- interface table and Fresnel events ------------------------------------------------------
N_SLAB_PACKETS = 200_000


def _fractions(st, n):
    return {k: st[k] / n for k in ("killed_surface", "escaped_t", "escaped_r", "absorbed")}


def _within(got, want, n, zmax=4.0):
    for k, w in want.items():
        sd = math.sqrt(max(w * (1 - w), 1.0 / n) / n)
        assert abs(got[k] - w) <= zmax * sd, (k, got[k], w)


@pytest.mark.parametrize("theta", [30.0, 45.0])
def test_clear_slab_fresnel_fractions(theta):
    sc = clear_slab(theta)
    st = O.run(sc, N_SLAB_PACKETS, seed=21, tally=False)["stats"]
    assert st["launched"] == N_SLAB_PACKETS and st["killed_bounce"] == 0
    _within(_fractions(st, N_SLAB_PACKETS), clear_slab_expected(theta), N_SLAB_PACKETS)


@pytest.mark.parametrize("theta", [30.0, 45.0])
def test_clear_slab_ray_geometry(theta):
    """Transmitted packets: emit -> refract (top) -> refract (bottom) -> escape; inside, Snell's
    angle asin(sin(theta)/1.5); after the slab, the entry direction again."""
    tr = O.run(clear_slab(theta), 200, seed=22, tally=False, trace_packets=200, trace_events=6)["trace"]
    th = math.radians(theta)
    d_in = np.array([math.sin(th), 0.0, -math.cos(th)])
    tt = math.asin(math.sin(th) / 1.5)
    d_slab = np.array([math.sin(tt), 0.0, -math.cos(tt)])
    seen = 0
    for row in tr:
        dirs, kinds = ray_directions(row)
        if list(kinds) != [1, 3, 3, 6]:
            assert list(kinds) in ([1, 4, 7], [1, 3, 4, 7]), kinds  # reflected at a face: killed
            continue
        seen += 1
        np.testing.assert_allclose(dirs[0], d_in, atol=1e-7)
        np.testing.assert_allclose(dirs[1], d_slab, atol=1e-7)
        np.testing.assert_allclose(dirs[2], d_in, atol=1e-7)
    assert seen > 150


def test_clear_slab_grazing_crossing_is_killed():
    """At 60 degrees the crossing step (min|d| + 2 eps along the ray) ends at most
    (min|d| + 2 eps) cos 60 - min|d| <= eps below the face, within eps: the next flight start
    kills the packet (src/inttau2.f90:52-56 after :112-116) whatever the Fresnel draw."""
    st = O.run(clear_slab(60.0), 50_000, seed=23, tally=False)["stats"]
    assert st["killed_surface"] == 50_000


def test_clear_slab_from_inside():
    """Inside the slab heading up: at 45 degrees, beyond the critical angle asin(1/1.5) = 41.8,
    every packet is totally reflected (R = 1) and killed; at 30 degrees the reflected fraction is
    the textbook R(30, 1.5 -> 1) and the rest leaves upwards."""
    st = O.run(clear_slab(45.0, inside=True), 50_000, seed=24, tally=False)["stats"]
    assert st["killed_surface"] == 50_000 and st["ev_fresnel"] >= 50_000
    R = fresnel_textbook(math.radians(30.0), 1.5, 1.0)
    st = O.run(clear_slab(30.0, inside=True), N_SLAB_PACKETS, seed=25, tally=False)["stats"]
    _within(_fractions(st, N_SLAB_PACKETS), {"killed_surface": R, "escaped_t": 1 - R}, N_SLAB_PACKETS)
