"""Reconstruction-quality metrics restated in numpy -- TEST INFRASTRUCTURE ONLY.

The checker for pmmc_amd/csrc/metrics.cpp.  Each function follows the
reference's definition (reference src/ paths):
  pearson               MetricsFramework::pearsonCorrelation (MetricsFramework.hpp:238-266)
  distance_correlation  MetricsFramework::distanceCorrelation (:268-286) with
                        BenchmarkRunner::runSingle's subsampling (BenchmarkRunner.hpp:314-331)
  scc                   MetricsFramework::structuralSimilarity / perDiagonalCorrelation (:419-479)
  rmsd                  Chromosome::calcRMSD -> fast_rmsd (Chromosome.hpp:613-634,
                        thirdparty/rmsd.cpp:395), here by Kabsch's SVD construction
                        (W. Kabsch, Acta Cryst. A32, 1976; reflection removed by the
                        sign of det), an independent method from the product's
                        quaternion eigenvalue.
Pinned by tests/test_metrics.py against the reference's own MetricsFramework
and fast_rmsd compiled in place (fixture tests/golden/metrics_ref.json) and
against closed forms.
"""
from __future__ import annotations

import numpy as np


def pearson(a, b) -> float:
    a = np.asarray(a, np.float32).astype(np.float64).reshape(-1)
    b = np.asarray(b, np.float32).astype(np.float64).reshape(-1)
    if a.size != b.size or a.size == 0:
        return 0.0
    da, db = a - a.mean(), b - b.mean()
    den = np.sqrt(np.dot(da, da) * np.dot(db, db))
    if den < 1e-12:
        return 0.0
    return float(np.float32(np.dot(da, db) / den))


def pair_distances(x, idx) -> np.ndarray:
    """FP32 lengths of FP32 differences (vector3::length, mtxlib.cpp:322) for
    the pairs i < j of idx, row by row."""
    x = np.asarray(x, np.float32).reshape(-1, 3)[idx]
    i, j = np.triu_indices(len(idx), 1)
    d = x[i] - x[j]
    s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return np.sqrt(s.astype(np.float64)).astype(np.float32)


def distance_correlation(x, y, step: int = 0) -> float:
    x = np.asarray(x, np.float32).reshape(-1, 3)
    y = np.asarray(y, np.float32).reshape(-1, 3)
    n = min(len(x), len(y))
    if n < 2:
        return 0.0
    if step <= 0:
        step = n // 500 if n > 500 else 1
    idx = np.arange(0, n, step)
    return pearson(pair_distances(x, idx), pair_distances(y, idx))


def distance_map(x) -> np.ndarray:
    x = np.asarray(x, np.float32).reshape(-1, 3)
    d = x[:, None, :] - x[None, :, :]
    s = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return np.sqrt(s.astype(np.float64)).astype(np.float32)


def scc(a, b, h: int = 5) -> float:
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    n = min(a.shape[0], b.shape[0])
    if n < 2:
        return 0.0
    wsum = wtot = 0.0
    for k in range(1, n):
        if n - k < 2:
            continue
        c = pearson(np.diagonal(a[:n, :n], k), np.diagonal(b[:n, :n], k))
        w = (n - k) * np.exp(-float(k * k) / (2.0 * h * h))
        wsum += w * c
        wtot += w
    if wtot < 1e-12:
        return 0.0
    return float(np.float32(wsum / wtot))


def rmsd(x, y) -> float:
    a = np.asarray(x, np.float32).astype(np.float64).reshape(-1, 3)
    b = np.asarray(y, np.float32).astype(np.float64).reshape(-1, 3)
    n = min(len(a), len(b))
    a, b = a[:n] - a[:n].mean(0), b[:n] - b[:n].mean(0)
    u, s, vt = np.linalg.svd(b.T @ a)
    d = np.sign(np.linalg.det(u @ vt))
    e = (a * a).sum() + (b * b).sum() - 2.0 * (s[0] + s[1] + d * s[2])
    return float(np.float32(np.sqrt(abs(e) / n)))
