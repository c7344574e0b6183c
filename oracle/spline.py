"""CPU restatement of the equidistant-model interpolation -- TEST
INFRASTRUCTURE ONLY (imported by tests/, never by pmmc_amd/).

HierarchicalChromosome::gpuHelper (HierarchicalChromosome.cu:1252-1449): the
CPU path (:1384-1446) and the kernel parallelFind3DSmoothPosition (:420-464)
compute, for output k at genomic position g = pos_start + k*resolution,
  g <  first_end            -> the first bead
  g >= last_start           -> the last bead
  else                      -> Catmull-Rom blend of beads i-2..i+1 for the
                               first i >= 1 with g < gp[i], t = (g - gp[i-1]) /
                               (gp[i] - gp[i-1]), mirrored end points
  no such i                 -> the last bead
in float, each product and sum rounded separately (the CPU path).

scan_positions restates the reference's loop literally (linear scan, scalar
float32 arithmetic); spline_positions is the vectorised form the tests use at
size.  PINNED BY tests/test_spline.py: the two agree bit for bit, t = 0
returns the bead exactly, a straight equally spaced chain maps to the line,
and the end and mirror rules on hand cases.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def count(pos_start, pos_stop, resolution):
    """Outputs the reference allocates: (stop - start + 1) / res + 1 (:1268)."""
    if resolution <= 0 or pos_stop < pos_start:
        return 0
    return (pos_stop - pos_start + 1) // resolution + 1


def _blend(t, p0, p1, p2, p3):
    t2 = F(t * t)
    t3 = F(t2 * t)
    b1 = F(F(0.5) * F(F(F(-t3) + F(F(2.0) * t2)) - t))
    b2 = F(F(0.5) * F(F(F(F(3.0) * t3) - F(F(5.0) * t2)) + F(2.0)))
    b3 = F(F(0.5) * F(F(F(F(-3.0) * t3) + F(F(4.0) * t2)) + t))
    b4 = F(F(0.5) * F(t3 - t2))
    return [F(F(F(F(p0[a] * b1) + F(p1[a] * b2)) + F(p2[a] * b3)) + F(p3[a] * b4)) for a in range(3)]


def scan_positions(gp, xyz, first_end, last_start, pos_start, resolution, n_out):
    """The reference CPU path, line by line (HierarchicalChromosome.cu:1384-1446)."""
    x = [[F(v) for v in row] for row in np.asarray(xyz, np.float32).reshape(-1, 3)]
    n = len(gp)
    out = np.zeros((n_out, 3), np.float32)
    for k in range(n_out):
        g = pos_start + resolution * k
        if g < first_end:
            out[k] = x[0]
            continue
        if g >= last_start:
            out[k] = x[n - 1]
            continue
        prev, found = -1, False
        for i in range(n):
            if g < gp[i] and prev >= 0:
                p1, p2 = x[prev], x[i]
                p0 = x[i - 2] if i > 1 else [F(F(F(2.0) * p1[a]) - p2[a]) for a in range(3)]
                p3 = x[i + 1] if i + 1 < n else [F(F(F(2.0) * p2[a]) - p1[a]) for a in range(3)]
                t = F(F(g - gp[prev]) / F(gp[i] - gp[prev]))
                out[k] = _blend(t, p0, p1, p2, p3)
                found = True
                break
            prev = i
        if not found:
            out[k] = x[prev]
    return out


def spline_positions(gp, xyz, first_end, last_start, pos_start, resolution, n_out):
    """Vectorised restatement (same per-element float32 operations)."""
    gp = np.asarray(gp, np.int64)
    x = np.asarray(xyz, np.float32).reshape(-1, 3)
    n = len(gp)
    g = pos_start + resolution * np.arange(n_out, dtype=np.int64)
    out = np.empty((n_out, 3), np.float32)
    first = g < first_end
    last = ~first & (g >= last_start)
    mid = ~first & ~last
    out[first] = x[0]
    out[last] = x[n - 1]
    i = np.maximum(np.searchsorted(gp, g, side="right"), 1)
    tail = mid & (i >= n)
    out[tail] = x[n - 1]
    m = mid & (i < n)
    ii = i[m]
    p1, p2 = x[ii - 1], x[ii]
    two = F(2.0)
    p0 = np.where((ii > 1)[:, None], x[np.maximum(ii - 2, 0)], two * p1 - p2)
    p3 = np.where((ii + 1 < n)[:, None], x[np.minimum(ii + 1, n - 1)], two * p2 - p1)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (g[m] - gp[ii - 1]).astype(np.float32) / (gp[ii] - gp[ii - 1]).astype(np.float32)
    t2 = t * t
    t3 = t2 * t
    b1 = F(0.5) * ((-t3 + F(2.0) * t2) - t)
    b2 = F(0.5) * ((F(3.0) * t3 - F(5.0) * t2) + F(2.0))
    b3 = F(0.5) * ((F(-3.0) * t3 + F(4.0) * t2) + t)
    b4 = F(0.5) * (t3 - t2)
    out[m] = ((p0 * b1[:, None] + p1 * b2[:, None]) + p2 * b3[:, None]) + p3 * b4[:, None]
    return out
