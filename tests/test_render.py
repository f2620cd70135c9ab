"""`render_sub` (src/sdfs/sdf_base.f90:302-346): the scene's minimum signed distance on the
render grid.  Pins: the oracle's restatement against the closed form of a sphere in a box,
including Fortran nint's half-away rounding of samples/2 (CPU); the GPU against the oracle on
every primitive, operator and modifier."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
from signedmcrt_amd.scene import Scene, cart_grid, source
from signedmcrt_amd.sdf import box, mono, sphere

M = mono(1.0, 0.1, 0.0, 1.0)


def test_oracle_render_closed_form():
    sc = Scene([sphere(0.5, M, 1), box((3.0, 3.0, 3.0), M, 2)], cart_grid(4, 4, 4, 1.0, 0.8, 1.2),
               source("point"))
    samples = (3, 4, 5)
    img = O.render(sc, samples)
    # ns = nint(samples / 2) = (2, 2, 3): 1.5 and 2.5 round away from zero
    ns = np.array([2, 2, 3])
    wid = np.array([1.0, 0.8, 1.2]) / ns
    want = np.empty(samples)
    for i in range(samples[0]):
        for j in range(samples[1]):
            for k in range(samples[2]):
                p = (np.array([i, j, k]) + 1 - ns) * wid
                q = np.abs(p) - 1.5
                dbox = np.linalg.norm(np.maximum(q, 0)) + min(q.max(), 0)
                want[i, j, k] = min(np.linalg.norm(p) - 0.5, dbox)
    np.testing.assert_allclose(img, want, atol=1e-12)
    # the voxel (2, 2, 3) sits on the origin: min(sphere -0.5, box -1.5)
    assert img[1, 1, 2] == pytest.approx(-1.5)


@pytest.mark.gpu
def test_render_matches_oracle(gpu):
    from signedmcrt_amd.engine import Engine
    from test_gpu_parity import sdf_zoo
    for k, s in enumerate(sdf_zoo()):
        sc = Scene([s, box((4.0, 4.0, 4.0), M, 2)], cart_grid(8, 8, 8, 1.5, 1.5, 1.5), source("point"))
        samples = (24, 20, 17)
        g = Engine(sc).render(samples)
        o = O.render(sc, samples)
        assert g.shape == o.shape == samples
        np.testing.assert_allclose(g, o, atol=2e-5 * (1 + np.abs(o)).max(), err_msg=str(k))
