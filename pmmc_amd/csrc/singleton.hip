// singleton.hip -- contact heatmap of a level's clusters from singleton PET
// records (LooperSolver::createSingletonHeatmap, LooperSolver.hpp:2517-2749).
//
// The O(records) part runs on the device: one pass bins both ends of every
// record (binary search over the chromosome's sorted breaks; a linear scan
// with the reference's first-match rule where they are not sorted), sums the
// scores as 64-bit integers -- in a per-workgroup LDS histogram when the whole
// N x N matrix fits, straight into the device matrix otherwise -- and keeps
// the counters and the chromosomes' extreme positions with one atomic per
// workgroup.  The O(N^2) remainder (long-range arcs, bin
// lengths, normalisation, diagonal) is the reference's float arithmetic, in
// its order, on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"

#pragma clang fp contract(off)

namespace pmmc {

namespace {

constexpr int kSingletonThreads = 256;
constexpr int kSingletonMaxChr = 64;             // per-workgroup min/max slots
constexpr int kSingletonLdsCells = 6144;         // 48 KB of 64-bit cells
constexpr int kSingletonLdsTable = 8192;         // 32 KB of breaks and tables
constexpr int32_t kLastBreak = 1000000000;       // LooperSolver.hpp:2536
constexpr int32_t kMaxPosition = 2000000000;     // :2554

struct SingletonArgs {
  int64_t n_records;
  const int32_t *chr1, *start1, *end1, *chr2, *start2, *end2, *score;
  int32_t n_chr, n;                  // chromosomes, bins
  const int32_t *breaks;             // all chromosomes' breaks, concatenated
  const int32_t *break_off;          // [n_chr]: first break of chromosome c
  const int32_t *bins;               // [n_chr]: bins of chromosome c
  const int32_t *bin0;               // [n_chr]: first bin index of chromosome c
  const int32_t *sorted;             // [n_chr]: breaks non-decreasing
  int32_t region_start, region_end;
  unsigned long long *cells;         // [n*n] score sums
  int32_t *cmin, *cmax;              // [n_chr]
  unsigned long long *counts;        // added, diagonal, bad_chr, unmatched
};

// first bin i with b[i] <= pos <= b[i+1], or -1 (LooperSolver.hpp:2652-2663)
__device__ __forceinline__ int find_bin(const int32_t *b, int nb, bool sorted, int32_t pos) {
  if (sorted) {
    // with non-decreasing breaks the first match is k - 1 for the first
    // k >= 1 with b[k] >= pos, provided b[k - 1] <= pos
    int lo = 1, hi = nb + 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (b[mid] >= pos) hi = mid;
      else lo = mid + 1;
    }
    if (lo > nb || b[lo - 1] > pos) return -1;
    return lo - 1;
  }
  for (int i = 0; i < nb; ++i)
    if (b[i] <= pos && pos <= b[i + 1]) return i;
  return -1;
}

// LDS: score sums in a per-workgroup histogram (the whole matrix fits);
// TBL: the breaks and per-chromosome tables copied to LDS (tbl_n ints: the
// bin searches' dependent reads then cost LDS, not L2, latency).
template <bool LDS, bool TBL>
__global__ __launch_bounds__(kSingletonThreads) void k_singleton_bin(SingletonArgs a, int tbl_n) {
  __shared__ unsigned long long hist[LDS ? kSingletonLdsCells : 1];
  __shared__ int32_t smin[kSingletonMaxChr], smax[kSingletonMaxChr];
  __shared__ unsigned long long scount[4];
  extern __shared__ int32_t stbl[];
  const int nn = a.n * a.n;
  if constexpr (LDS)
    for (int k = threadIdx.x; k < nn; k += blockDim.x) hist[k] = 0;
  for (int c = threadIdx.x; c < a.n_chr; c += blockDim.x) {
    smin[c] = kMaxPosition;
    smax[c] = 0;
  }
  if (threadIdx.x < 4) scount[threadIdx.x] = 0;
  // the tables are contiguous: breaks, break_off, bins, bin0, sorted
  const int32_t *breaks = a.breaks, *break_off = a.break_off, *bins = a.bins, *bin0 = a.bin0,
                *sorted = a.sorted;
  if constexpr (TBL) {
    for (int k = threadIdx.x; k < tbl_n; k += blockDim.x) stbl[k] = a.breaks[k];
    breaks = stbl;
    break_off = stbl + (a.break_off - a.breaks);
    bins = break_off + a.n_chr;
    bin0 = bins + a.n_chr;
    sorted = bin0 + a.n_chr;
  }
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  // counters per thread, summed per workgroup at the end: one global atomic
  // per counter and workgroup (per-record atomics on four addresses serialise)
  uint32_t n_added = 0, n_diag = 0, n_bad = 0, n_unmatched = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n_records;
       i += stride) {
    bool bad = false, diag = false, added = false, unmatched = false;
    {
      const int c1 = a.chr1[i], c2 = a.chr2[i];
      if (c1 < 0 || c1 >= a.n_chr || c2 < 0 || c2 >= a.n_chr) {
        bad = true;
      } else {
        int32_t s1 = a.start1[i];
        const int32_t e1 = a.end1[i];
        int32_t s2 = a.start2[i];
        const int32_t e2 = a.end2[i];
        const bool in_region = a.region_end <= 0 ||
                               (s1 >= a.region_start && s1 <= a.region_end && e2 >= a.region_start &&
                                e2 <= a.region_end);
        if (in_region) {
          s1 = s1 < 0 ? 0 : s1;
          s2 = s2 < 0 ? 0 : s2;
          const int b1 = find_bin(breaks + break_off[c1], bins[c1], sorted[c1] != 0, s1);
          const int b2 = find_bin(breaks + break_off[c2], bins[c2], sorted[c2] != 0, s2);
          if (b1 < 0 || b2 < 0) {
            unmatched = true;
          } else {
            const int st = bin0[c1] + b1, en = bin0[c2] + b2;
            if (st == en) {
              diag = true;
            } else {
              added = true;
              // the extremes change rarely: atomics only when improving
              if (s1 < smin[c1]) atomicMin(&smin[c1], s1);
              if (e1 > smax[c1]) atomicMax(&smax[c1], e1);
              if (s2 < smin[c2]) atomicMin(&smin[c2], s2);
              if (e2 > smax[c2]) atomicMax(&smax[c2], e2);
              const unsigned long long sc =
                  static_cast<unsigned long long>(static_cast<int64_t>(a.score[i]));
              // [st][en] and [en][st] receive the same sums: one add to the
              // upper-triangle cell, mirrored on the host
              const int lo = st < en ? st : en, hi = st < en ? en : st;
              if constexpr (LDS)
                atomicAdd(&hist[lo * a.n + hi], sc);
              else
                atomicAdd(&a.cells[static_cast<size_t>(lo) * a.n + hi], sc);
            }
          }
        }
      }
    }
    n_added += added;
    n_diag += diag;
    n_bad += bad;
    n_unmatched += unmatched;
  }
  if (n_added) atomicAdd(&scount[0], static_cast<unsigned long long>(n_added));
  if (n_diag) atomicAdd(&scount[1], static_cast<unsigned long long>(n_diag));
  if (n_bad) atomicAdd(&scount[2], static_cast<unsigned long long>(n_bad));
  if (n_unmatched) atomicAdd(&scount[3], static_cast<unsigned long long>(n_unmatched));
  __syncthreads();
  if (threadIdx.x < 4 && scount[threadIdx.x]) atomicAdd(&a.counts[threadIdx.x], scount[threadIdx.x]);
  if constexpr (LDS)
    for (int k = threadIdx.x; k < nn; k += blockDim.x)
      if (hist[k]) atomicAdd(&a.cells[k], hist[k]);
  for (int c = threadIdx.x; c < a.n_chr; c += blockDim.x) {
    if (smin[c] != kMaxPosition) atomicMin(&a.cmin[c], smin[c]);
    if (smax[c] != 0) atomicMax(&a.cmax[c], smax[c]);
  }
}

__global__ void k_singleton_clear(SingletonArgs a) {
  const size_t nn = static_cast<size_t>(a.n) * a.n;
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < nn;
       k += static_cast<size_t>(gridDim.x) * blockDim.x)
    a.cells[k] = 0;
  if (blockIdx.x == 0) {
    for (int c = threadIdx.x; c < a.n_chr; c += blockDim.x) {
      a.cmin[c] = kMaxPosition;
      a.cmax[c] = 0;
    }
    if (threadIdx.x < 4) a.counts[threadIdx.x] = 0;
  }
}

int first_bin_host(const std::vector<int32_t> &b, int off, int nb, int32_t pos) {
  for (int i = 0; i < nb; ++i)
    if (b[off + i] <= pos && pos <= b[off + i + 1]) return i;
  return -1;
}

// LooperSolver::getGenomicLengthExcludingCentromere (LooperSolver.hpp:5916-5935)
int32_t length_excluding_centromere(bool has, int32_t cs, int32_t ce, int32_t st, int32_t end) {
  if (!has) return end - st + 1;
  if (cs > end || ce < st) return end - st + 1;
  if (cs < st) {
    if (ce < end) return end - ce + 1;
    return 0;
  }
  if (ce < end) return end - st - (ce - cs + 1) + 1;
  return cs - st;
}

}  // namespace

}  // namespace pmmc

using namespace pmmc;

extern "C" pmmc_status pmmc_singleton_heatmap(pmmc_ctx *ctx, const pmmc_singleton_input *in,
                                              float *heat, pmmc_singleton_counts *counts) {
  if (!ctx || !in || !heat) return fail(ctx, PMMC_ERR_INVALID, "singleton: null argument");
  const int n_chr = in->n_chr;
  if (n_chr < 1 || n_chr > kSingletonMaxChr || !in->cluster_offsets || in->n_records < 0 ||
      in->n_long < 0 || in->diag < 0)
    return fail(ctx, PMMC_ERR_INVALID, "singleton: 1..64 chromosomes, non-negative counts");
  if (in->n_records > 0 && (!in->chr1 || !in->start1 || !in->end1 || !in->chr2 || !in->start2 ||
                            !in->end2 || !in->score))
    return fail(ctx, PMMC_ERR_INVALID, "singleton: null record array");
  if (in->n_long > 0 && (!in->long_chr || !in->long_start || !in->long_end || !in->long_score))
    return fail(ctx, PMMC_ERR_INVALID, "singleton: null long-arc array");
  // breaks (LooperSolver.hpp:2527-2540)
  std::vector<int32_t> breaks, break_off(n_chr), bins(n_chr), bin0(n_chr), sorted(n_chr);
  int n = 0;
  for (int c = 0; c < n_chr; ++c) {
    const int k0 = in->cluster_offsets[c], k1 = in->cluster_offsets[c + 1];
    if (k1 < k0 || (k1 > k0 && (!in->cluster_start || !in->cluster_end)))
      return fail(ctx, PMMC_ERR_INVALID, "singleton: bad cluster offsets");
    break_off[c] = static_cast<int32_t>(breaks.size());
    breaks.push_back(0);
    for (int k = k0; k + 1 < k1; ++k)
      breaks.push_back((in->cluster_end[k] + in->cluster_start[k + 1]) / 2);
    breaks.push_back(kLastBreak);
    bins[c] = static_cast<int32_t>(breaks.size()) - break_off[c] - 1;
    bin0[c] = n;
    n += bins[c];
    sorted[c] = 1;
    for (int k = break_off[c]; k + 1 < static_cast<int>(breaks.size()); ++k)
      if (breaks[k + 1] < breaks[k]) sorted[c] = 0;
  }
  if (static_cast<int64_t>(n) * n > (int64_t(1) << 31))
    return fail(ctx, PMMC_ERR_INVALID, "singleton: more than 46,340 bins");
  const size_t nn = static_cast<size_t>(n) * n;
  // device buffers: cells, 4 counters, min/max, tables, and the records staged
  // in chunks of up to 2^24
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(in->n_records, int64_t(1) << 24));
  const size_t rec_bytes = static_cast<size_t>(chunk) * 4;
  const size_t tbl = breaks.size() + 4 * static_cast<size_t>(n_chr);
  auto rounded = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t bytes = rounded(nn * 8) + rounded(64) + 2 * rounded(4 * n_chr) + rounded(4 * tbl) +
                       7 * rounded(rec_bytes);
  pmmc_status ps = arena_reserve(ctx, bytes);
  if (ps != PMMC_OK) return ps;
  char *p = ctx->arena;
  auto take = [&](size_t b) {
    char *q = p;
    p += rounded(b);
    return q;
  };
  SingletonArgs a;
  a.cells = reinterpret_cast<unsigned long long *>(take(nn * 8));
  a.counts = reinterpret_cast<unsigned long long *>(take(64));
  a.cmin = reinterpret_cast<int32_t *>(take(4 * n_chr));
  a.cmax = reinterpret_cast<int32_t *>(take(4 * n_chr));
  int32_t *d_tbl = reinterpret_cast<int32_t *>(take(4 * tbl));
  int32_t *d_rec[7];
  for (int k = 0; k < 7; ++k) d_rec[k] = reinterpret_cast<int32_t *>(take(rec_bytes));
  std::vector<int32_t> h_tbl(breaks);
  h_tbl.insert(h_tbl.end(), break_off.begin(), break_off.end());
  h_tbl.insert(h_tbl.end(), bins.begin(), bins.end());
  h_tbl.insert(h_tbl.end(), bin0.begin(), bin0.end());
  h_tbl.insert(h_tbl.end(), sorted.begin(), sorted.end());
  a.breaks = d_tbl;
  a.break_off = d_tbl + breaks.size();
  a.bins = a.break_off + n_chr;
  a.bin0 = a.bins + n_chr;
  a.sorted = a.bin0 + n_chr;
  a.n_chr = n_chr;
  a.n = n;
  a.n_records = in->n_records;
  a.region_start = in->region_start;
  a.region_end = in->region_end;
  hipStream_t st = ctx->stream;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_tbl, h_tbl.data(), 4 * tbl, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_singleton_clear, dim3(static_cast<unsigned>(std::min<size_t>(1024, (nn + 255) / 256 + 1))),
                     dim3(256), 0, st, a);
  PMMC_HIP_TRY(ctx, hipGetLastError());
  unsigned long long h_counts[4] = {0, 0, 0, 0};
  double kernel_ms = 0.0;
  const int32_t *h_rec[7] = {in->chr1, in->start1, in->end1, in->chr2, in->start2, in->end2, in->score};
  const bool lds = nn <= static_cast<size_t>(kSingletonLdsCells);
  const int tbl_n = static_cast<int>(tbl);
  const bool tbl_lds = tbl <= static_cast<size_t>(kSingletonLdsTable);
  const size_t dyn = tbl_lds ? tbl * 4 : 0;
  for (int64_t r0 = 0; r0 < in->n_records; r0 += chunk) {
    const int64_t m = std::min<int64_t>(chunk, in->n_records - r0);
    for (int k = 0; k < 7; ++k)
      PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_rec[k], h_rec[k] + r0, static_cast<size_t>(m) * 4,
                                       hipMemcpyHostToDevice, st));
    a.n_records = m;
    a.chr1 = d_rec[0];
    a.start1 = d_rec[1];
    a.end1 = d_rec[2];
    a.chr2 = d_rec[3];
    a.start2 = d_rec[4];
    a.end2 = d_rec[5];
    a.score = d_rec[6];
    const int64_t per_group = (lds ? 64 : 4) * static_cast<int64_t>(kSingletonThreads);
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(
        1, std::min<int64_t>(static_cast<int64_t>(ctx->cu_count) * 8, (m + per_group - 1) / per_group)));
    PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    if (lds && tbl_lds)
      hipLaunchKernelGGL((k_singleton_bin<true, true>), dim3(grid), dim3(kSingletonThreads), dyn, st, a, tbl_n);
    else if (lds)
      hipLaunchKernelGGL((k_singleton_bin<true, false>), dim3(grid), dim3(kSingletonThreads), 0, st, a, tbl_n);
    else if (tbl_lds)
      hipLaunchKernelGGL((k_singleton_bin<false, true>), dim3(grid), dim3(kSingletonThreads), dyn, st, a, tbl_n);
    else
      hipLaunchKernelGGL((k_singleton_bin<false, false>), dim3(grid), dim3(kSingletonThreads), 0, st, a, tbl_n);
    PMMC_HIP_TRY(ctx, hipGetLastError());
    PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    // the chunk's host arrays stay in use until its copies ran
    PMMC_HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    PMMC_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    kernel_ms += ms;
  }
  std::vector<unsigned long long> cells(nn);
  std::vector<int32_t> cmin(n_chr), cmax(n_chr);
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(cells.data(), a.cells, nn * 8, hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(h_counts, a.counts, sizeof(h_counts), hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(cmin.data(), a.cmin, 4 * n_chr, hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(cmax.data(), a.cmax, 4 * n_chr, hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(st));
  if (h_counts[3] != 0)
    return fail(ctx, PMMC_ERR_INVALID, "singleton: a record position lies outside every bin");
  // ---- host: the reference's float steps, in its order
  // the device summed each pair into its upper-triangle cell; both cells of
  // the reference's symmetric sums get it
  std::vector<float> h(nn, 0.0f);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const float v = static_cast<float>(static_cast<int64_t>(cells[static_cast<size_t>(i) * n + j]));
      h[static_cast<size_t>(i) * n + j] = v;
      h[static_cast<size_t>(j) * n + i] = v;
    }
  // long-range arcs, chromosome by chromosome (LooperSolver.hpp:2701-2737)
  for (int c = 0; c < n_chr; ++c)
    for (int k = 0; k < in->n_long; ++k) {
      if (in->long_chr[k] != c) continue;
      const int i = first_bin_host(breaks, break_off[c], bins[c], in->long_start[k]);
      const int j = first_bin_host(breaks, break_off[c], bins[c], in->long_end[k]);
      if (i < 0 || j < 0) return fail(ctx, PMMC_ERR_INVALID, "singleton: a long arc lies outside every bin");
      if (i == j) continue;
      const float val = static_cast<float>(in->long_scale *
                                           std::pow(static_cast<double>(in->long_score[k]),
                                                    static_cast<double>(in->long_power)));
      float &cell = h[static_cast<size_t>(bin0[c] + i) * n + bin0[c] + j];
      cell += val;  // twice into [st][end], as written (:2735-2736)
      cell += val;
    }
  // extreme positions replace the outer breaks; bin lengths (:2740-2767)
  std::vector<float> len(n);
  for (int c = 0; c < n_chr; ++c) {
    int32_t *b = breaks.data() + break_off[c];
    b[0] = cmin[c];
    b[bins[c]] = cmax[c];
    const bool has = in->centromere && in->centromere[3 * c] != 0;
    const int32_t cs = has ? in->centromere[3 * c + 1] : 0, ce = has ? in->centromere[3 * c + 2] : 0;
    for (int i = 0; i < bins[c]; ++i) {
      const int32_t d = length_excluding_centromere(has, cs, ce, b[i], b[i + 1]);
      len[bin0[c] + i] = static_cast<float>(d) / 1e6f;
    }
  }
  // normalise by bin sizes, mirrored (:2773-2779); clear the band
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const float sc = h[static_cast<size_t>(i) * n + j] / (len[i] * len[j]);
      h[static_cast<size_t>(i) * n + j] = sc;
      h[static_cast<size_t>(j) * n + i] = sc;
    }
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < in->diag && i + q < n; ++q) {
      h[static_cast<size_t>(i) * n + i + q] = 0.0f;
      h[static_cast<size_t>(i + q) * n + i] = 0.0f;
    }
  std::memcpy(heat, h.data(), nn * sizeof(float));
  if (counts) {
    counts->added = static_cast<int64_t>(h_counts[0]);
    counts->diagonal = static_cast<int64_t>(h_counts[1]);
    counts->bad_chr = static_cast<int64_t>(h_counts[2]);
    counts->kernel_ms = kernel_ms;
  }
  return PMMC_OK;
}
