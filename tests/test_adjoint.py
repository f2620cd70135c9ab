"""Perturbation-MC adjoint (no reference counterpart): d(R_d)/d(mu_a, mu_s) and
d(sum_v w_v jmean_v / N)/d(mu_a, mu_s) per material.

Pins: the oracle's FP64 estimator against common-random-number central differences of the
oracle itself (CPU); the GPU kernel against the oracle's gradients (independent samples, batch
standard errors) and against GPU common-random-number differences; the torch.autograd binding
against the direct call.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
from signedmcrt_amd.scene import Scene, cart_grid, source
from signedmcrt_amd.sdf import box, mono, sphere


def slab_scene(mus=20.0, mua=1.0):
    arr = [box((1.0, 1.0, 0.2), mono(mus, mua, 0.8, 1.0), 1), box((2, 2, 2), mono(0, 0, 0, 1.0), 2)]
    return Scene(arr, cart_grid(50, 50, 50, 0.5, 0.5, 0.25), source("pencil", (0, 0, 0.2), (0, 0, -1)))


def two_layer_scene(mus=(20.0, 30.0), mua=(1.0, 0.5)):
    """A sphere of a second material inside the slab: two scattering materials + air."""
    top = mono(mus[0], mua[0], 0.8, 1.0)
    inc = mono(mus[1], mua[1], 0.9, 1.0)
    arr = [sphere(0.06, inc, 1), box((1.0, 1.0, 0.2), top, 2), box((2, 2, 2), mono(0, 0, 0, 1.0), 3)]
    return Scene(arr, cart_grid(50, 50, 50, 0.5, 0.5, 0.25), source("pencil", (0, 0, 0.2), (0, 0, -1)))


def _fd_oracle(make, which, mat, h, n, B, seed):
    """CRN central differences of (R, W) per batch."""
    out = []
    for b in range(B):
        v = []
        for s in (+1, -1):
            sc = make()
            m = sc.materials[mat]
            if which == "mua":
                m.mua += s * h
            else:
                m.mus += s * h
            sc = Scene(sc.array, sc.grid, sc.source)
            r = O.run(sc, n, seed=seed, first_packet=b * n)
            v.append((r["stats"]["escaped_r"] / n, r["jmean"].sum() / n))
        out.append(((v[0][0] - v[1][0]) / (2 * h), (v[0][1] - v[1][1]) / (2 * h)))
    return np.array(out)


def _grad_batches(run, B):
    g = [run(b) for b in range(B)]
    keys = ("d_r_dmua", "d_r_dmus", "d_w_dmua", "d_w_dmus")
    return {k: np.array([x[k] for x in g]) for k in keys}


def _z(a, b):
    """z of the difference of two batch means (arrays of per-batch values)."""
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return (a.mean() - b.mean()) / max(se, 1e-12)


@pytest.mark.parametrize("which,h", [("mua", 0.1), ("mus", 1.0)])
def test_oracle_score_gradient_matches_crn_differences(which, h):
    B, n = 6, 40000
    g = _grad_batches(lambda b: O.run(slab_scene(), n, seed=5, first_packet=b * n, grad=True), B)
    fd = _fd_oracle(slab_scene, which, 0, h, n, B, seed=5)
    zr = _z(g[f"d_r_d{which}"][:, 0], fd[:, 0])
    zw = _z(g[f"d_w_d{which}"][:, 0], fd[:, 1])
    assert abs(zr) < 4 and abs(zw) < 4, (zr, zw, g[f"d_r_d{which}"].mean(0), fd.mean(0))


def test_weighted_tally_identity():
    """w = 1: the gradient run's W equals the plain tally sum / N of the same packets."""
    sc = slab_scene()
    g = O.run(sc, 20000, seed=3, grad=True)
    r = O.run(sc, 20000, seed=3)
    assert g["w_tally"] == pytest.approx(r["jmean"].sum() / 20000, rel=1e-12)


# ---------------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle(gpu):
    from signedmcrt_amd.engine import Engine
    sc = two_layer_scene()
    eng = Engine(sc)
    B, n, no = 8, 400_000, 60_000
    gg = _grad_batches(lambda b: eng.grad(n, seed=11, first_packet=b * n).__dict__, B)
    go = _grad_batches(lambda b: O.run(sc, no, seed=12, first_packet=b * no, grad=True), B)
    for k in gg:
        for m in range(sc.n_materials):
            z = _z(gg[k][:, m], go[k][:, m])
            assert abs(z) < 4.0, (k, m, z, gg[k][:, m].mean(), go[k][:, m].mean())


@pytest.mark.gpu
@pytest.mark.parametrize("which,h", [("mua", 0.1), ("mus", 1.5)])
def test_gpu_adjoint_matches_gpu_crn_differences(gpu, which, h):
    from signedmcrt_amd.engine import Engine
    sc = two_layer_scene()
    eng = Engine(sc)
    B, n = 8, 1_000_000
    g = _grad_batches(lambda b: eng.grad(n, seed=21, first_packet=b * n).__dict__, B)
    mus0 = np.array([m.mus for m in sc.materials])
    mua0 = np.array([m.mua for m in sc.materials])
    for mat in (0, 1):
        fd = []
        for b in range(B):
            v = []
            for s in (+1, -1):
                mus, mua = mus0.copy(), mua0.copy()
                (mua if which == "mua" else mus)[mat] += s * h
                eng.set_materials(mus=mus, mua=mua)
                r = eng.run(n, seed=21, first_packet=b * n)
                v.append((r.stats["escaped_r"] / n, r.jmean.sum() / n))
            fd.append(((v[0][0] - v[1][0]) / (2 * h), (v[0][1] - v[1][1]) / (2 * h)))
        eng.set_materials(mus=mus0, mua=mua0)
        fd = np.array(fd)
        zr = _z(g[f"d_r_d{which}"][:, mat], fd[:, 0])
        zw = _z(g[f"d_w_d{which}"][:, mat], fd[:, 1])
        assert abs(zr) < 4 and abs(zw) < 4, (mat, zr, zw)


@pytest.mark.gpu
def test_autograd_function_matches_direct_gradients(gpu):
    import torch
    from signedmcrt_amd.autograd import MCRTFluence
    from signedmcrt_amd.engine import Engine
    sc = two_layer_scene()
    eng = Engine(sc)
    mua = torch.tensor([m.mua for m in sc.materials], dtype=torch.float64, requires_grad=True)
    mus = torch.tensor([m.mus for m in sc.materials], dtype=torch.float64, requires_grad=True)
    n = 400_000
    phi, rd = MCRTFluence.apply(mua, mus, eng, n, 31)
    loss = phi.sum() + 2.0 * rd
    loss.backward()
    g = eng.grad(n, seed=31)
    np.testing.assert_allclose(mua.grad.numpy(), g.d_w_dmua + 2.0 * g.d_r_dmua, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(mus.grad.numpy(), g.d_w_dmus + 2.0 * g.d_r_dmus, rtol=1e-9, atol=1e-12)
