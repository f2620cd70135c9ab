"""Inverse fit of absorption coefficients from a fluence map with the perturbation-MC adjoint (no
reference counterpart).

    hist = fit_mua(engine, phi_meas, materials=[0, ..., 15], mua0=0.5, n_packets=2e7)

Loss L(mu) = sum_c (Phi_c(mu) - Meas_c)^2 / sum_c Meas_c^2 over blocks c of `block`^3 voxels
(block = 1: every voxel), Phi_c the block sum of phi = raw path-length tally / N.  Each
iteration runs the forward loop at the current mu_a, then one adjoint launch over the same
packets with voxel weights w_v = dL/dphi_v = 2 (Phi_c - Meas_c) / norm for v in c, which returns
dL/dmu_a for every selected material at once (`Engine.grad`); Adam steps on log mu_a.  Block
sums average the per-voxel Monte Carlo noise that otherwise sets the loss floor.  With
`lr_final` the step size decays from `lr` to `lr_final` on a cosine schedule, and `average_last`
> 0 returns the mean of the last iterates in log mu_a (Polyak averaging): both damp the
fluctuation of the noisy gradient steps around the optimum.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np


@dataclass
class FitHistory:
    mua: List[np.ndarray] = field(default_factory=list)   # fitted mu_a after each iteration
    loss: List[float] = field(default_factory=list)
    grad: List[np.ndarray] = field(default_factory=list)  # dL/dmu_a at each iteration
    averaged: Optional[np.ndarray] = None  # mean of the last iterates (log space), if asked for

    @property
    def final(self) -> np.ndarray:
        return self.averaged if self.averaged is not None else self.mua[-1]


def _blocks(a: np.ndarray, f: int) -> np.ndarray:
    nx, ny, nz = a.shape
    return a.reshape(nx // f, f, ny // f, f, nz // f, f).sum(axis=(1, 3, 5))


def fit_mua(engine, phi_meas: np.ndarray, materials: Sequence[int], mua0, n_packets: int,
            iters: int = 30, lr: float = 0.1, seed: int = 1000, betas=(0.9, 0.999), block: int = 1,
            lr_final: Optional[float] = None, average_last: int = 0,
            callback: Optional[Callable[[int, np.ndarray, float], None]] = None) -> FitHistory:
    """Fit mu_a of `materials` (indices into engine.scene.materials) to `phi_meas` (same shape as
    the grid, fluence per launched packet).  Other materials keep their properties.  `block`
    must divide the grid's dimensions."""
    if average_last < 0 or average_last > iters:
        raise ValueError("average_last must lie in [0, iters]")
    if any(d % block for d in phi_meas.shape):
        raise ValueError(f"block {block} does not divide the grid {phi_meas.shape}")
    mats = list(materials)
    base = np.array([m.mua for m in engine.scene.materials], dtype=float)
    mus = np.array([m.mus for m in engine.scene.materials], dtype=float)
    theta = np.log(np.broadcast_to(np.asarray(mua0, dtype=float), (len(mats),)).copy())
    m1 = np.zeros_like(theta)
    m2 = np.zeros_like(theta)
    meas_b = _blocks(phi_meas.astype(float), block)
    norm = float(np.sum(meas_b ** 2)) or 1.0
    n = int(n_packets)
    hist = FitHistory()
    for it in range(iters):
        mua = base.copy()
        mua[mats] = np.exp(theta)
        engine.set_materials(mus=mus, mua=mua)
        r = engine.run(n, seed=seed + it)
        resid = _blocks(r.jmean / n, block) - meas_b
        loss = float(np.sum(resid ** 2) / norm)
        w = (2.0 / norm) * np.repeat(np.repeat(np.repeat(resid, block, 0), block, 1), block, 2)
        g = engine.grad(n, seed=seed + it, weights=w, materials=mats)
        dl = np.asarray(g.d_w_dmua, dtype=float)[mats]
        gt = dl * np.exp(theta)  # chain rule to log mu_a
        m1 = betas[0] * m1 + (1 - betas[0]) * gt
        m2 = betas[1] * m2 + (1 - betas[1]) * gt * gt
        mh = m1 / (1 - betas[0] ** (it + 1))
        vh = m2 / (1 - betas[1] ** (it + 1))
        step = lr
        if lr_final is not None and iters > 1:
            step = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * it / (iters - 1)))
        theta = theta - step * mh / (np.sqrt(vh) + 1e-12)
        hist.loss.append(loss)
        hist.grad.append(dl)
        hist.mua.append(np.exp(theta))
        if callback:
            callback(it, np.exp(theta), loss)
    if average_last:
        hist.averaged = np.exp(np.mean(np.log(np.array(hist.mua[-average_last:])), axis=0))
        theta = np.log(hist.averaged)
    mua = base.copy()
    mua[mats] = np.exp(theta)
    engine.set_materials(mus=mus, mua=mua)
    return hist
