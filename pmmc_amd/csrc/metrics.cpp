// metrics.cpp -- reconstruction-quality metrics used to compare annealing
// results with a ground-truth structure and with each other: distance-map
// Pearson correlation, stratum-adjusted correlation (SCC) of two maps, and
// the RMSD after optimal superposition.
//
// Definitions follow the reference's metrics (reference src/ paths):
//   pmmc_pearson               MetricsFramework::pearsonCorrelation
//                              (MetricsFramework.hpp:238-266) and the identical
//                              BenchmarkRunner::calcDistanceCorrelation
//                              (BenchmarkRunner.hpp:340-368): FP64 sums in
//                              input order, 0 for empty / mismatched input or
//                              a degenerate denominator (< 1e-12)
//   pmmc_distance_correlation  MetricsFramework::distanceCorrelation
//                              (MetricsFramework.hpp:268-286) with the
//                              subsampling rule of BenchmarkRunner::runSingle
//                              (BenchmarkRunner.hpp:314-331): FP32 lengths of
//                              FP32 differences (vector3::length, mtxlib.cpp:322)
//   pmmc_distance_map          the FP32 pairwise length matrix of a structure
//   pmmc_scc                   MetricsFramework::structuralSimilarity and
//                              perDiagonalCorrelation (MetricsFramework.hpp:419-479)
//   pmmc_rmsd                  Chromosome::calcRMSD (Chromosome.hpp:613-634),
//                              i.e. fast_rmsd (thirdparty/rmsd.cpp:395): the
//                              least RMSD over proper rotations and translations.
//                              Computed here with Horn's unit-quaternion method
//                              (B.K.P. Horn, "Closed-form solution of absolute
//                              orientation using unit quaternions", JOSA A 4(4),
//                              1987): RMSD^2 = (|A|^2 + |B|^2 - 2 lambda_max) / n,
//                              lambda_max the largest eigenvalue of Horn's 4x4
//                              key matrix, found by cyclic Jacobi rotations.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/pmmc_levels.h"

namespace {

float pearson_impl(const float *a, const float *b, int64_t n) {
  if (n <= 0) return 0.0f;
  double sa = 0.0, sb = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    sa += a[i];
    sb += b[i];
  }
  const double ma = sa / static_cast<double>(n), mb = sb / static_cast<double>(n);
  double cov = 0.0, va = 0.0, vb = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double da = a[i] - ma, db = b[i] - mb;
    cov += da * db;
    va += da * da;
    vb += db * db;
  }
  const double den = std::sqrt(va * vb);
  if (den < 1e-12) return 0.0f;
  return static_cast<float>(cov / den);
}

// FP32 length of the FP32 difference of two beads (float sum of squares,
// then the square root rounded to float).
inline float bead_distance(const float *x, int i, int j) {
  const float dx = x[3 * i] - x[3 * j];
  const float dy = x[3 * i + 1] - x[3 * j + 1];
  const float dz = x[3 * i + 2] - x[3 * j + 2];
  const float s = dx * dx + dy * dy + dz * dz;
  return static_cast<float>(std::sqrt(static_cast<double>(s)));
}

// Largest eigenvalue of a symmetric 4x4 matrix (cyclic Jacobi).
double max_eigenvalue_sym4(double M[4][4]) {
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < 4; ++p) {
      diag += M[p][p] * M[p][p];
      for (int q = p + 1; q < 4; ++q) off += M[p][q] * M[p][q];
    }
    if (off <= 1e-30 * (diag + 1e-300)) break;
    for (int p = 0; p < 3; ++p) {
      for (int q = p + 1; q < 4; ++q) {
        if (M[p][q] == 0.0) continue;
        const double theta = (M[q][q] - M[p][p]) / (2.0 * M[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {  // columns p, q
          const double mkp = M[k][p], mkq = M[k][q];
          M[k][p] = c * mkp - s * mkq;
          M[k][q] = s * mkp + c * mkq;
        }
        for (int k = 0; k < 4; ++k) {  // rows p, q
          const double mpk = M[p][k], mqk = M[q][k];
          M[p][k] = c * mpk - s * mqk;
          M[q][k] = s * mpk + c * mqk;
        }
      }
    }
  }
  double best = M[0][0];
  for (int p = 1; p < 4; ++p) best = std::fmax(best, M[p][p]);
  return best;
}

}  // namespace

extern "C" float pmmc_pearson(const float *a, const float *b, int64_t n) {
  if (!a || !b) return 0.0f;
  return pearson_impl(a, b, n);
}

extern "C" float pmmc_distance_correlation(const float *xyz_a, const float *xyz_b, int32_t n,
                                           int32_t step) {
  if (!xyz_a || !xyz_b || n < 2) return 0.0f;
  if (step <= 0) step = n > 500 ? n / 500 : 1;
  std::vector<float> da, db;
  for (int i = 0; i < n; i += step)
    for (int j = i + step; j < n; j += step) {
      da.push_back(bead_distance(xyz_a, i, j));
      db.push_back(bead_distance(xyz_b, i, j));
    }
  return pearson_impl(da.data(), db.data(), static_cast<int64_t>(da.size()));
}

extern "C" pmmc_status pmmc_distance_map(const float *xyz, int32_t n, float *out) {
  if (!xyz || !out || n < 0) return PMMC_ERR_INVALID;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[static_cast<int64_t>(i) * n + j] = i == j ? 0.0f : bead_distance(xyz, i, j);
  return PMMC_OK;
}

extern "C" float pmmc_scc(const float *a, const float *b, int32_t n, int32_t h_param) {
  if (!a || !b || n < 2) return 0.0f;
  double wsum = 0.0, wtot = 0.0;
  std::vector<float> va, vb;
  for (int k = 1; k < n; ++k) {
    va.clear();
    vb.clear();
    for (int i = 0; i + k < n; ++i) {
      va.push_back(a[static_cast<int64_t>(i) * n + i + k]);
      vb.push_back(b[static_cast<int64_t>(i) * n + i + k]);
    }
    if (va.size() < 2) continue;
    const float corr = pearson_impl(va.data(), vb.data(), static_cast<int64_t>(va.size()));
    if (std::isnan(corr)) continue;
    const int pairs = n - k;
    const double w = pairs * std::exp(-static_cast<double>(k * k) /
                                      (2.0 * h_param * h_param));
    wsum += w * corr;
    wtot += w;
  }
  if (wtot < 1e-12) return 0.0f;
  return static_cast<float>(wsum / wtot);
}

extern "C" float pmmc_rmsd(const float *xyz_a, const float *xyz_b, int32_t n) {
  if (!xyz_a || !xyz_b || n <= 0) return 0.0f;
  double ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      ca[k] += xyz_a[3 * i + k];
      cb[k] += xyz_b[3 * i + k];
    }
  for (int k = 0; k < 3; ++k) {
    ca[k] /= n;
    cb[k] /= n;
  }
  double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double ga = 0.0, gb = 0.0;
  for (int i = 0; i < n; ++i) {
    double pa[3], pb[3];
    for (int k = 0; k < 3; ++k) {
      pa[k] = xyz_a[3 * i + k] - ca[k];
      pb[k] = xyz_b[3 * i + k] - cb[k];
      ga += pa[k] * pa[k];
      gb += pb[k] * pb[k];
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) S[r][c] += pb[r] * pa[c];
  }
  const double xx = S[0][0], xy = S[0][1], xz = S[0][2];
  const double yx = S[1][0], yy = S[1][1], yz = S[1][2];
  const double zx = S[2][0], zy = S[2][1], zz = S[2][2];
  double N[4][4] = {
      {xx + yy + zz, yz - zy, zx - xz, xy - yx},
      {yz - zy, xx - yy - zz, xy + yx, zx + xz},
      {zx - xz, xy + yx, -xx + yy - zz, yz + zy},
      {xy - yx, zx + xz, yz + zy, -xx - yy + zz},
  };
  const double lmax = max_eigenvalue_sym4(N);
  const double msd = std::fabs(ga + gb - 2.0 * lmax) / n;
  return static_cast<float>(std::sqrt(msd));
}
