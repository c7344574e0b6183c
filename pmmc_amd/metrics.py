"""Reconstruction-quality metrics (include/pmmc_levels.h, host C++ in
libpmmc_hip.so): distance-map Pearson, SCC and RMSD after superposition, as
the reference's MetricsFramework / BenchmarkRunner define them."""
from __future__ import annotations

import numpy as np

from . import _abi as A


def _xyz(x):
    return np.ascontiguousarray(x, np.float32).reshape(-1, 3)


def pearson(a, b) -> float:
    """MetricsFramework::pearsonCorrelation (MetricsFramework.hpp:238-266)."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    b = np.ascontiguousarray(b, np.float32).reshape(-1)
    if a.size != b.size:
        return 0.0
    return float(A.load().pmmc_pearson(a.ctypes.data_as(A.f32p), b.ctypes.data_as(A.f32p), a.size))


def distance_correlation(x, y, step: int = 0) -> float:
    """Pearson of pairwise distances (MetricsFramework::distanceCorrelation,
    MetricsFramework.hpp:268-286); step 0 = BenchmarkRunner's n/500 rule."""
    x, y = _xyz(x), _xyz(y)
    n = min(len(x), len(y))
    return float(A.load().pmmc_distance_correlation(x.ctypes.data_as(A.f32p),
                                                    y.ctypes.data_as(A.f32p), n, step))


def distance_map(x) -> np.ndarray:
    x = _xyz(x)
    out = np.empty((len(x), len(x)), np.float32)
    st = A.load().pmmc_distance_map(x.ctypes.data_as(A.f32p), len(x), out.ctypes.data_as(A.f32p))
    if st != 0:
        raise ValueError("pmmc_distance_map failed")
    return out


def scc(a, b, h: int = 5) -> float:
    """MetricsFramework::structuralSimilarity (MetricsFramework.hpp:419-456)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    n = min(a.shape[0], b.shape[0])
    a = np.ascontiguousarray(a[:n, :n])
    b = np.ascontiguousarray(b[:n, :n])
    return float(A.load().pmmc_scc(a.ctypes.data_as(A.f32p), b.ctypes.data_as(A.f32p), n, h))


def rmsd(x, y) -> float:
    """Chromosome::calcRMSD / fast_rmsd: least RMSD over rotations and translations."""
    x, y = _xyz(x), _xyz(y)
    n = min(len(x), len(y))
    return float(A.load().pmmc_rmsd(x.ctypes.data_as(A.f32p), y.ctypes.data_as(A.f32p), n))


def compare(recon, truth, h: int = 5) -> dict:
    """The reconstruction-vs-ground-truth figures the reference's benchmark
    reports (BenchmarkRunner.hpp:300-337) plus the SCC of the two distance maps."""
    return {"pearson": distance_correlation(recon, truth),
            "scc": scc(distance_map(recon), distance_map(truth), h),
            "rmsd": rmsd(recon, truth)}
