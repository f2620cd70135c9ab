"""Statistical comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np


def two_sample_z(k1: int, n1: int, k2: int, n2: int) -> float:
    """z of the difference of two binomial proportions (pooled variance)."""
    p = (k1 + k2) / (n1 + n2)
    var = p * (1 - p) * (1.0 / n1 + 1.0 / n2)
    if var <= 0:
        return 0.0 if k1 * n2 == k2 * n1 else np.inf
    return (k1 / n1 - k2 / n2) / np.sqrt(var)


def coarsen(grid: np.ndarray, f: int) -> np.ndarray:
    """Sum f x f x f blocks of voxels (the dimensions must be multiples of f)."""
    nx, ny, nz = grid.shape
    return grid.reshape(nx // f, f, ny // f, f, nz // f, f).sum(axis=(1, 3, 5))


def batch_mean_se(batches):
    """Per-cell mean and standard error over a list of equally sized batches."""
    a = np.stack(batches)
    m = a.mean(axis=0)
    se = a.std(axis=0, ddof=1) / np.sqrt(len(batches))
    return m, se


def field_z(m1, se1, m2, se2, rel_floor: float = 0.01):
    """z-scores of two estimated fields on cells holding >= rel_floor of the peak."""
    mask = np.maximum(m1, m2) >= rel_floor * max(m1.max(), m2.max())
    den = np.sqrt(se1[mask] ** 2 + se2[mask] ** 2)
    ok = den > 0
    return (m1[mask][ok] - m2[mask][ok]) / den[ok]
