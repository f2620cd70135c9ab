"""Scenes with closed-form answers, shared by the oracle pins (CPU) and the GPU parity tests."""
from __future__ import annotations

import math

import numpy as np

from signedmcrt_amd.scene import Scene, cart_grid, source
from signedmcrt_amd.sdf import box, mono

N_SLAB = 1.5
SLAB_THICK = 0.2


def fresnel_textbook(theta_i: float, n1: float, n2: float) -> float:
    """Unpolarised Fresnel reflectance from the angle forms Rs = sin^2(ti-tt)/sin^2(ti+tt),
    Rp = tan^2(ti-tt)/tan^2(ti+tt) with Snell's law; 1 beyond the critical angle."""
    s = n1 / n2 * math.sin(theta_i)
    if s >= 1.0:
        return 1.0
    tt = math.asin(s)
    rs = (math.sin(theta_i - tt) / math.sin(theta_i + tt)) ** 2
    rp = (math.tan(theta_i - tt) / math.tan(theta_i + tt)) ** 2
    return 0.5 * (rs + rp)


def clear_slab(theta_deg: float, inside: bool = False) -> Scene:
    """A clear (mu_s = mu_a = 0) n = 1.5 slab, 4 x 4 x 0.2 cm, in air, hit by a collimated beam at
    theta from the normal in the xz plane.  From outside the beam starts 0.1 cm above the top face
    heading down; `inside` starts it at the slab's mid-plane heading up.  Nothing scatters or
    absorbs, so every packet follows the same ray and only the Fresnel draws differ."""
    th = math.radians(theta_deg)
    arr = [box((4.0, 4.0, SLAB_THICK), mono(0.0, 0.0, 0.0, N_SLAB), 1),
           box((6.0, 6.0, 6.0), mono(0.0, 0.0, 0.0, 1.0), 2)]
    if inside:
        src = source("pencil", (0.0, 0.0, 0.0), (math.sin(th), 0.0, math.cos(th)))
    else:
        src = source("pencil", (-0.2 * math.tan(th), 0.0, SLAB_THICK / 2 + 0.1), (math.sin(th), 0.0, -math.cos(th)))
    return Scene(arr, cart_grid(16, 16, 16, 2.5, 2.5, 1.0), src)


def clear_slab_expected(theta_deg: float):
    """Fractions for `clear_slab` from outside under the reference's semantics: a Fresnel
    reflection steps the packet back to within eps of the surface and the next flight start kills
    it (src/inttau2.f90:58-61,170-175), so killed = R + (1 - R) R and transmitted = (1 - R)^2
    (the internal angle at the bottom face equals the refraction angle at the top, and the
    reflectance is the same both ways)."""
    R = fresnel_textbook(math.radians(theta_deg), 1.0, N_SLAB)
    return {"killed_surface": R + (1 - R) * R, "escaped_t": (1 - R) ** 2, "escaped_r": 0.0, "absorbed": 0.0}


def ray_directions(trace_row):
    """Unit directions between successive recorded events of one traced packet."""
    k = trace_row[:, 3]
    pts = trace_row[k > 0, :3].astype(np.float64)
    d = np.diff(pts, axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return d / np.linalg.norm(d, axis=1)[:, None], k[k > 0].astype(int)
