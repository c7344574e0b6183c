// stress.hip -- gradient-descent stress refinement of anchor positions
// (LooperSolver::stressRefineAnchors, LooperSolver.hpp:5638-5915), one
// workgroup per problem, every iteration inside the kernel.
//
// Per iteration, at positions X:
//   backbone  spring i (beads i, i+1, target b_i > 1e-6):
//               e = (|X_i - X_i+1| - b)/b, E += k e^2 (k = k_stretch for
//               e >= 0, else k_squeeze), gradient 2k(d - b)/(b^2 d) (X_i - X_i+1)
//   pairs     pair p (i, j, target t): E += ((|X_i - X_j| - t)/t)^2, gradient
//               2(d - t)/(t^2 d) (X_i - X_j) (distances clamped below at 1e-9)
//   bending   window i = 1..n-2: E += k_angular theta_i^3, theta the angle
//               between X_i - X_i-1 and X_i+1 - X_i; gradient by forward
//               differences (step 1e-3) over the window's 3 beads x 3 axes
//   step      if the energy rose, halve the step (stop below 1e-6, before
//               moving); X -= step * G; stop when the energy fell by less
//               than 1e-7 max(1, E) after iteration 21.
//
// Layout: a problem runs on one or more workgroups (as many as its chain
// needs at one bead per thread, within what the device keeps resident); each
// thread owns beads and computes a bead's whole gradient itself -- its two
// springs, its pairs (a per-bead list of (pair, other bead, target) entries
// in pair order) and its three bending windows -- so an iteration is one
// pass, a barrier, the decision (taken by every workgroup from the same
// per-workgroup partial sums), the update of its own beads and a barrier.
// Several workgroups of one problem meet at a grid barrier (cooperative
// launch, so all are resident; every wait is bounded and a timeout ends the
// kernel with an error instead of a hang).  A spring, pair or window shared
// by several beads is evaluated by each of them with the same operations,
// hence the same bits.
//
// Results are the reference's bit for bit:
//  * every float operation rounds where the reference's does (no fused
//    multiply-adds in this file, IEEE division and square root, the
//    reference's acosf by acosf_fdlibm);
//  * each bead's gradient is summed in the order the reference's loops
//    scatter into it: its springs (i-1 then i), its pairs in pair order, its
//    bending windows i-1, i, i+1;
//  * the energy totals only steer the two decisions (halve the step, stop).
//    The reference sums its terms in order, one at a time; here the terms are
//    summed by a fixed tree and each decision is certified against the
//    rounding-error bound of both orders, |seq - tree| <= (nt + 16) 2^-52
//    sum|terms|; a decision inside the bound (or any non-finite value) falls
//    back to the in-order sums of this and the previous iteration's terms
//    (both kept), so the decision is always the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "acosf_fdlibm.hpp"
#include "ctx.hpp"
#include "pmmc_hip.h"

#pragma clang fp contract(off)

namespace pmmc {

namespace {

constexpr int kStressThreads = 256;
constexpr int kStressWaves = kStressThreads / 64;
// polls of the grid barrier before it gives up (each poll sleeps ~64 cycles:
// about 0.1-0.2 s, far beyond any wait of a resident grid)
constexpr unsigned kSpinLimit = 1u << 22;

struct StressDev {
  int n, np, nt, max_iter;
  int in_order, n_wg;  // workgroups of this problem
  float lr, ks, kq, ka;
  float *x;            // [3n] positions, updated in place
  const float *bb;     // [n-1] spring targets
  const int *goff;     // [n+1] per-bead entry lists
  const int4 *ent;     // {pair, other bead, target bits, 1 if the bead is the pair's second end}
  float *g;            // [3n] gradient (each thread's own beads)
  double *terms;       // [2][nt]: springs (n-1), pairs (np), windows (n-2)
  double *part;        // [2][n_wg][4] per-workgroup partial sums
  unsigned *bar;       // [2] grid barrier: arrivals, completed generation
  struct StressOut *out;
};

struct StressOut {
  pmmc_stress_stats s;
  int buf;             // terms buffer of the last evaluated iteration
};

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ V3 ld3(const float *p, int i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// vector3::length (mtxlib.cpp:322): sqrt of (x*x + y*y) + z*z in float
__device__ __forceinline__ float len(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// bending energy of a window; false when an arm is shorter than 1e-9
__device__ __forceinline__ bool bend(V3 a, V3 b, V3 c, float ka, float &E) {
  const V3 v1 = sub(b, a), v2 = sub(c, b);
  const float n1 = len(v1), n2 = len(v2);
  if (n1 < 1e-9f || n2 < 1e-9f) return false;
  float cs = dot(v1, v2) / (n1 * n2);
  if (cs > 1.0f) cs = 1.0f;
  if (cs < -1.0f) cs = -1.0f;
  const float th = acosf_fdlibm(cs);
  E = ka * th * th * th;
  return true;
}

// gradient vector of a spring or pair d = X_i - X_j with target t:
// 2k(|d| - t)/(t^2 |d|) d, with e = (|d| - t)/t and k = k_st (e >= 0) or k_sq
__device__ __forceinline__ V3 spring_grad(V3 d, float t, float k_st, float k_sq, float &k, float &e) {
  float dl = len(d);
  if (dl < 1e-9f) dl = 1e-9f;
  e = (dl - t) / t;
  k = e >= 0.0f ? k_st : k_sq;
  const float gm = 2.0f * k * (dl - t) / (t * t * dl);
  return V3{d.x * gm, d.y * gm, d.z * gm};
}

// One in-order sum of m terms by a whole wavefront: 64 consecutive terms are
// loaded at once, then added one at a time in index order (every lane keeps
// the same running sum).  Padding adds +0.0, which leaves a sum that started
// at +0.0 unchanged.
__device__ double wave_seq_sum(const double *t, int m) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int base = 0; base < m; base += 64) {
    const double v = base + lane < m ? t[base + lane] : 0.0;
    const int lo = __double2loint(v), hi = __double2hiint(v);
#pragma unroll
    for (int j = 0; j < 64; ++j)
      acc += __hiloint2double(__builtin_amdgcn_readlane(hi, j), __builtin_amdgcn_readlane(lo, j));
  }
  return acc;
}

// the reference's sums of one iteration's terms: e_back, e_hic, e_ang and
// (e_back + e_hic) + e_ang (called by a whole wavefront)
__device__ double seq_sums(const double *t, int n1, int np, int n3, double &eb, double &eh,
                           double &ea) {
  eb = wave_seq_sum(t, n1);
  eh = wave_seq_sum(t + n1, np);
  ea = wave_seq_sum(t + n1 + np, n3);
  return eb + eh + ea;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ double std_max1(double t) { return 1.0 < t ? t : 1.0; }  // std::max(1.0, t)

// Barrier over the problem's workgroups.  Every thread's global writes are
// released at device scope before the workgroup meets; thread 0 counts the
// arrival (arrivals only grow: generation g is complete at n_wg * g) and the
// last one publishes the generation; after the wait every thread acquires.
// false (and *err set) when the wait exceeded kSpinLimit polls.
__device__ bool grid_barrier(const StressDev &P, unsigned &gen, int *err) {
  if (P.n_wg == 1) {
    __syncthreads();
    return true;
  }
  __shared__ int s_ok;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned target = gen + 1u;
    int ok = 1;
    const unsigned arrived =
        __hip_atomic_fetch_add(&P.bar[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (arrived == static_cast<unsigned>(P.n_wg) * target) {
      __hip_atomic_store(&P.bar[1], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(&P.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins == kSpinLimit) {
          ok = 0;
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    gen = target;
    s_ok = ok;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return s_ok != 0;
}

__global__ __launch_bounds__(kStressThreads) void k_stress(const StressDev *__restrict__ probs,
                                                           const int2 *__restrict__ blocks,
                                                           int *err) {
  const int2 bm = blocks[blockIdx.x];  // {problem, workgroup within it}
  const StressDev &P = probs[bm.x];
  const int wg = bm.y;
  const int tid = threadIdx.x;
  const int n = P.n, np = P.np, nt = P.nt, n_wg = P.n_wg;
  const int n1 = n - 1, n3 = n - 2;
  const float ks = P.ks, kq = P.kq, ka = P.ka;
  const float h = 1e-3f;
  const float *__restrict__ bb = P.bb;
  const int *__restrict__ goff = P.goff;
  const int4 *__restrict__ ent = P.ent;
  float *__restrict__ G = P.g;
  float *X = P.x;
  __shared__ double red[4][kStressWaves];
  __shared__ int s_update, s_done;
  __shared__ float s_lr;
  // decision state (thread 0 of every workgroup, identical in all)
  double prevT = 1e300, prevB = 0.0, prev_seq = 0.0;
  bool have_prev_seq = false;
  float lr = P.lr;
  int iters = 0, stop = PMMC_STRESS_MAX_ITER, exact_count = 0, last = 0;
  unsigned gen = 0;

  for (int it = 0; it < P.max_iter; ++it) {
    double *__restrict__ T = P.terms + static_cast<size_t>(it & 1) * nt;
    double pb = 0.0, ph = 0.0, pa = 0.0, pabs = 0.0;
    for (int k = wg * kStressThreads + tid; k < n; k += n_wg * kStressThreads) {
      const V3 xk = ld3(X, k);
      float gx = 0.0f, gy = 0.0f, gz = 0.0f;
      float kk, e;
      // spring k-1 (this bead is its second end), then spring k (first end)
      if (k >= 1) {
        const float t = bb[k - 1];
        if (!(t <= 1e-6f)) {
          const V3 gv = spring_grad(sub(ld3(X, k - 1), xk), t, ks, kq, kk, e);
          gx -= gv.x;
          gy -= gv.y;
          gz -= gv.z;
        }
      }
      if (k < n1) {
        const float t = bb[k];
        double term = 0.0;
        if (!(t <= 1e-6f)) {
          const V3 gv = spring_grad(sub(xk, ld3(X, k + 1)), t, ks, kq, kk, e);
          gx += gv.x;
          gy += gv.y;
          gz += gv.z;
          term = static_cast<double>(kk) * e * e;
        }
        T[k] = term;
        pb += term;
        pabs += fabs(term);
      }
      // pairs, in pair order (weight 1)
      const int e1 = goff[k + 1];
      for (int q = goff[k]; q < e1; ++q) {
        const int4 en = ent[q];
        const V3 xo = ld3(X, en.y);
        const float t = __int_as_float(en.z);
        if (en.w) {
          const V3 gv = spring_grad(sub(xo, xk), t, 1.0f, 1.0f, kk, e);
          gx -= gv.x;
          gy -= gv.y;
          gz -= gv.z;
        } else {
          const V3 gv = spring_grad(sub(xk, xo), t, 1.0f, 1.0f, kk, e);
          gx += gv.x;
          gy += gv.y;
          gz += gv.z;
          const double term = static_cast<double>(1.0f) * e * e;
          T[n1 + en.x] = term;
          ph += term;
          pabs += fabs(term);
        }
      }
      // bending windows k-1, k, k+1; this bead is bead 2, 1, 0 of them
      if (ka > 0.0f) {
#pragma unroll
        for (int dw = -1; dw <= 1; ++dw) {
          const int w = k + dw;
          if (w < 1 || w > n - 2) continue;
          const V3 a = dw == 1 ? xk : ld3(X, w - 1);
          const V3 b = dw == 0 ? xk : ld3(X, w);
          const V3 c = dw == -1 ? xk : ld3(X, w + 1);
          float E0;
          const bool ok = bend(a, b, c, ka, E0);
          if (dw == 0) {
            const double term = ok ? static_cast<double>(E0) : 0.0;
            T[n1 + np + w - 1] = term;
            pa += term;
            pabs += fabs(term);
          }
          if (!ok) continue;
          constexpr int kBead[3] = {2, 1, 0};  // this bead's place in window k + dw
          const int bi = kBead[dw + 1];
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            V3 q3[3] = {a, b, c};
            float *qp = ax == 0 ? &q3[bi].x : (ax == 1 ? &q3[bi].y : &q3[bi].z);
            *qp = *qp + h;
            float Ep;
            if (!bend(q3[0], q3[1], q3[2], ka, Ep)) continue;
            const float dE = (Ep - E0) / h;
            if (ax == 0) gx += dE;
            if (ax == 1) gy += dE;
            if (ax == 2) gz += dE;
          }
        }
      } else if (k >= 1 && k <= n - 2) {
        T[n1 + np + k - 1] = 0.0;
      }
      G[3 * k] = gx;
      G[3 * k + 1] = gy;
      G[3 * k + 2] = gz;
    }
    // per-workgroup partial sums (fixed order), then over the workgroups
    pb = wave_sum(pb);
    ph = wave_sum(ph);
    pa = wave_sum(pa);
    pabs = wave_sum(pabs);
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) {
      red[0][wv] = pb;
      red[1][wv] = ph;
      red[2][wv] = pa;
      red[3][wv] = pabs;
    }
    __syncthreads();
    double *part = P.part + static_cast<size_t>(it & 1) * n_wg * 4;
    if (tid < 4) {
      double v = 0.0;
      for (int w = 0; w < kStressWaves; ++w) v += red[tid][w];
      part[wg * 4 + tid] = v;
    }
    if (!grid_barrier(P, gen, err)) return;
    // sums over the workgroups' partials: thread t takes workgroups t,
    // t + 256, ..., then a fixed tree -- the same order in every workgroup
    {
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      for (int w = tid; w < n_wg; w += kStressThreads) {
        const double2 lo = *reinterpret_cast<const double2 *>(part + w * 4);
        const double2 hi = *reinterpret_cast<const double2 *>(part + w * 4 + 2);
        v[0] += lo.x;
        v[1] += lo.y;
        v[2] += hi.x;
        v[3] += hi.y;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = wave_sum(v[j]);
      __syncthreads();  // red[] of the partial sums above is no longer read
      if (lane == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[j][wv] = v[j];
      __syncthreads();
    }
    if (wv == 0) {  // the decision: every lane of wavefront 0 computes it alike
      double tb = 0.0, th = 0.0, ta = 0.0, sabs = 0.0;
      for (int w = 0; w < kStressWaves; ++w) {
        tb += red[0][w];
        th += red[1][w];
        ta += red[2][w];
        sabs += red[3][w];
      }
      const double Tt = tb + th + ta;
      const double B = static_cast<double>(nt + 16) * 0x1p-52 * sabs * 1.01;
      bool exact = P.in_order != 0 || !isfinite(Tt) || !isfinite(B);
      if (!exact && it > 0) {
        if (!(fabs(Tt - prevT) > B + prevB)) exact = true;
        if (it > 20) {
          const double f = prevT - Tt - 1e-7 * std_max1(Tt);
          const double m = prevB + B * (1.0 + 1e-7) + 8.0 * 0x1p-52 * (fabs(prevT) + fabs(Tt));
          if (!(fabs(f) > m)) exact = true;
        }
      }
      double total = Tt, prev = prevT, eb, eh, ea;
      if (exact) {
        ++exact_count;
        total = seq_sums(T, n1, np, n3, eb, eh, ea);
        if (it > 0) {
          const double *Tp = P.terms + static_cast<size_t>((it + 1) & 1) * nt;
          prev = have_prev_seq ? prev_seq : seq_sums(Tp, n1, np, n3, eb, eh, ea);
        }
      }
      last = it;
      int update = 1, done = 0;
      if (total > prev && it > 0) {
        lr *= 0.5f;
        if (lr < 1e-6f) {
          update = 0;
          done = 1;
          stop = PMMC_STRESS_LR_FLOOR;
        }
      }
      if (update) {
        ++iters;
        if (prev - total < 1e-7 * std_max1(total) && it > 20) {
          done = 1;
          stop = PMMC_STRESS_CONVERGED;
        }
      }
      prevT = Tt;
      prevB = B;
      prev_seq = total;
      have_prev_seq = exact;
      if (lane == 0) {
        s_update = update;
        s_done = done;
        s_lr = lr;
      }
    }
    __syncthreads();
    if (s_update) {
      const float l = s_lr;
      for (int k = wg * kStressThreads + tid; k < n; k += n_wg * kStressThreads) {
        X[3 * k] -= l * G[3 * k];
        X[3 * k + 1] -= l * G[3 * k + 1];
        X[3 * k + 2] -= l * G[3 * k + 2];
      }
    }
    const bool done = s_done != 0;
    if (!grid_barrier(P, gen, err)) return;
    if (done) break;
  }
  // statistics; the energies of the last evaluated positions are summed in
  // order by the host from that iteration's terms (buffer `buf`)
  if (wg == 0 && tid == 0) {
    StressOut S{};
    S.s.iterations = iters;
    S.s.stop = P.max_iter > 0 ? stop : PMMC_STRESS_NOTHING;
    S.s.lr = lr;
    S.s.exact_sums = exact_count;
    S.buf = last & 1;
    *P.out = S;
  }
}

// splitmix64 finaliser: a per-input term of the acosf checksum
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_acosf_sum(uint32_t first, uint32_t count,
                                                   unsigned long long *out) {
  uint64_t acc = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint32_t b = first + i;
    const float a = acosf_fdlibm(f32_of(b));
    acc += mix64((static_cast<uint64_t>(b) << 32) | bits_of(a));
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, static_cast<unsigned long long>(acc));
}

}  // namespace

}  // namespace pmmc

using namespace pmmc;

extern "C" pmmc_status pmmc_probe_acosf(pmmc_ctx *ctx, uint32_t first_bits, uint32_t count,
                                        uint64_t *sum) {
  if (!ctx || !sum) return PMMC_ERR_INVALID;
  if (static_cast<uint64_t>(first_bits) + count > 0x100000000ull)
    return fail(ctx, PMMC_ERR_INVALID, "probe_acosf: range past 2^32");
  pmmc_status ps = arena_reserve(ctx, 4096);
  if (ps != PMMC_OK) return ps;
  unsigned long long *d = reinterpret_cast<unsigned long long *>(ctx->arena);
  hipStream_t s = ctx->stream;
  PMMC_HIP_TRY(ctx, hipMemsetAsync(d, 0, sizeof(*d), s));
  if (count > 0) {
    const uint32_t blocks = std::min<uint32_t>((count + 255) / 256, 8192);
    hipLaunchKernelGGL(k_acosf_sum, dim3(blocks), dim3(256), 0, s, first_bits, count, d);
    PMMC_HIP_TRY(ctx, hipGetLastError());
  }
  unsigned long long h = 0;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(s));
  *sum = h;
  return PMMC_OK;
}

extern "C" void pmmc_stress_opts_default(pmmc_stress_opts *o) {
  if (!o) return;
  // stressRefineIterations / stressRefineLR defaults (Settings.hpp:408-409)
  o->max_iter = 0;
  o->lr = 0.05f;
  o->in_order_sums = 0;
}

extern "C" pmmc_status pmmc_stress_refine(pmmc_ctx *ctx, int32_t count,
                                          const pmmc_stress_problem *probs,
                                          const pmmc_stress_opts *opts, float *const *pos_out,
                                          pmmc_stress_stats *stats) {
  if (!ctx) return PMMC_ERR_INVALID;
  if (count < 0 || (count > 0 && (!probs || !opts || !pos_out)))
    return fail(ctx, PMMC_ERR_INVALID, "stress: null argument");
  // validate every problem before writing any output
  for (int32_t c = 0; c < count; ++c) {
    const pmmc_stress_problem &q = probs[c];
    if (q.n < 0 || (q.n > 0 && (!q.pos || !pos_out[c])) || q.n_pairs < 0 ||
        (q.n_pairs > 0 && (!q.pair_i || !q.pair_j || !q.pair_target)) || (q.n > 1 && !q.backbone))
      return fail(ctx, PMMC_ERR_INVALID, "stress: bad sizes or null arrays");
    for (int32_t p = 0; p < q.n_pairs; ++p)
      if (q.pair_i[p] < 0 || q.pair_i[p] >= q.n || q.pair_j[p] < 0 || q.pair_j[p] >= q.n ||
          q.pair_i[p] == q.pair_j[p])
        return fail(ctx, PMMC_ERR_INVALID, "stress: pair index out of range or i == j");
  }
  // problems with nothing to do keep their positions (LooperSolver.hpp:5640-5641:
  // max_iter <= 0 or fewer than 2 beads)
  std::vector<int> work;
  for (int32_t c = 0; c < count; ++c) {
    const pmmc_stress_problem &q = probs[c];
    if (q.n > 0 && pos_out[c] != q.pos) std::copy(q.pos, q.pos + 3 * static_cast<size_t>(q.n), pos_out[c]);
    if (stats) {
      stats[c] = pmmc_stress_stats{};
      stats[c].stop = PMMC_STRESS_NOTHING;
      stats[c].lr = opts->lr;
    }
    if (opts->max_iter > 0 && q.n >= 2) work.push_back(c);
  }
  if (work.empty()) return PMMC_OK;

  // workgroups per problem: one bead per thread where the device keeps the
  // whole grid resident, else `bpt` beads per thread
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stress, kStressThreads, 0) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  const size_t capacity = static_cast<size_t>(per_cu) * static_cast<size_t>(ctx->cu_count > 0 ? ctx->cu_count : 1);
  size_t bpt = 1, total_wg = 0;
  auto wgs_of = [&](int n) { return static_cast<int>((n + kStressThreads * bpt - 1) / (kStressThreads * bpt)); };
  for (;;) {
    total_wg = 0;
    for (int c : work) total_wg += wgs_of(probs[c].n);
    if (total_wg <= capacity || bpt >= (1u << 20)) break;
    bpt *= 2;
  }
  // per-bead entry lists in pair order: {pair, other bead, target, second end}
  std::vector<std::vector<int>> goff(work.size());
  std::vector<std::vector<int4>> ents(work.size());
  Sizer z;
  for (size_t w = 0; w < work.size(); ++w) {
    const pmmc_stress_problem &q = probs[work[w]];
    const int n = q.n, np = q.n_pairs;
    std::vector<int> &off = goff[w];
    std::vector<int4> &en = ents[w];
    off.assign(n + 1, 0);
    for (int p = 0; p < np; ++p) {
      ++off[q.pair_i[p] + 1];
      ++off[q.pair_j[p] + 1];
    }
    for (int k = 0; k < n; ++k) off[k + 1] += off[k];
    en.resize(2 * static_cast<size_t>(np));
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int p = 0; p < np; ++p) {
      int tb;
      std::memcpy(&tb, &q.pair_target[p], 4);
      en[fill[q.pair_i[p]]++] = int4{p, q.pair_j[p], tb, 0};
      en[fill[q.pair_j[p]]++] = int4{p, q.pair_i[p], tb, 1};
    }
    const size_t nt = static_cast<size_t>(n - 1) + np + (n - 2);
    z.add<float>(3 * n);                             // x
    z.add<float>(n);                                 // bb
    z.add<int>(n + 1);                               // goff
    z.add<int4>(2 * static_cast<size_t>(np) + 1);    // entries
    z.add<float>(3 * n);                             // g
    z.add<double>(2 * nt);                           // terms
    z.add<double>(8 * static_cast<size_t>(wgs_of(n)));  // partial sums
  }
  z.add<StressDev>(work.size());
  z.add<StressOut>(work.size());
  z.add<int2>(total_wg);
  z.add<unsigned>(2 * work.size() + 1);              // barriers, error flag
  pmmc_status ps = arena_reserve(ctx, z.top + 4096);
  if (ps != PMMC_OK) return ps;
  Carver cv{ctx->arena, 0, ctx->arena_cap};
  hipStream_t s = ctx->stream;
  std::vector<StressDev> hd(work.size());
  std::vector<int2> blk;
  blk.reserve(total_wg);
  for (size_t w = 0; w < work.size(); ++w) {
    const pmmc_stress_problem &q = probs[work[w]];
    const int n = q.n, np = q.n_pairs;
    StressDev &d = hd[w];
    d.n = n;
    d.np = np;
    d.nt = (n - 1) + np + (n - 2);
    d.max_iter = opts->max_iter;
    d.in_order = opts->in_order_sums;
    d.n_wg = wgs_of(n);
    d.lr = opts->lr;
    d.ks = q.k_stretch;
    d.kq = q.k_squeeze;
    d.ka = q.k_angular;
    d.x = cv.take<float>(3 * n);
    float *bb = cv.take<float>(n);
    int *go = cv.take<int>(n + 1);
    int4 *en = cv.take<int4>(2 * static_cast<size_t>(np) + 1);
    d.bb = bb;
    d.goff = go;
    d.ent = en;
    d.g = cv.take<float>(3 * n);
    d.terms = cv.take<double>(2 * static_cast<size_t>(d.nt));
    d.part = cv.take<double>(8 * static_cast<size_t>(d.n_wg));
    for (int g = 0; g < d.n_wg; ++g) blk.push_back(int2{static_cast<int>(w), g});
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(d.x, q.pos, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(bb, q.backbone, sizeof(float) * (n - 1), hipMemcpyHostToDevice, s));
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(go, goff[w].data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (np > 0)
      PMMC_HIP_TRY(ctx, hipMemcpyAsync(en, ents[w].data(), sizeof(int4) * 2 * np, hipMemcpyHostToDevice, s));
  }
  StressDev *dprobs = cv.take<StressDev>(work.size());
  StressOut *dstats = cv.take<StressOut>(work.size());
  int2 *dblk = cv.take<int2>(total_wg);
  unsigned *dbar = cv.take<unsigned>(2 * work.size() + 1);
  int *derr = reinterpret_cast<int *>(dbar + 2 * work.size());
  bool grid_sync = false;
  for (size_t w = 0; w < work.size(); ++w) {
    hd[w].out = dstats + w;
    hd[w].bar = dbar + 2 * w;
    grid_sync = grid_sync || hd[w].n_wg > 1;
  }
  PMMC_HIP_TRY(ctx, hipMemsetAsync(dbar, 0, sizeof(unsigned) * (2 * work.size() + 1), s));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(dprobs, hd.data(), sizeof(StressDev) * hd.size(), hipMemcpyHostToDevice, s));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(dblk, blk.data(), sizeof(int2) * blk.size(), hipMemcpyHostToDevice, s));
  if (grid_sync) {
    // workgroups of one problem wait for each other: the cooperative launch
    // keeps the whole grid resident (or refuses to start)
    void *args[] = {&dprobs, &dblk, &derr};
    PMMC_HIP_TRY(ctx, hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_stress),
                                                 dim3(static_cast<unsigned>(total_wg)),
                                                 dim3(kStressThreads), args, 0, s));
  } else {
    hipLaunchKernelGGL(k_stress, dim3(static_cast<unsigned>(total_wg)), dim3(kStressThreads), 0, s,
                       dprobs, dblk, derr);
    PMMC_HIP_TRY(ctx, hipGetLastError());
  }
  std::vector<StressOut> hs(work.size());
  int herr = 0;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(hs.data(), dstats, sizeof(StressOut) * hs.size(), hipMemcpyDeviceToHost, s));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, s));
  for (size_t w = 0; w < work.size(); ++w)
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(pos_out[work[w]], hd[w].x, sizeof(float) * 3 * hd[w].n,
                                     hipMemcpyDeviceToHost, s));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(s));
  if (herr) return fail(ctx, PMMC_ERR_HIP, "stress: a grid barrier wait timed out");
  if (stats) {
    // the reference's in-order sums of the last evaluated iteration's terms
    std::vector<double> t;
    for (size_t w = 0; w < work.size(); ++w) {
      const StressDev &d = hd[w];
      pmmc_stress_stats S = hs[w].s;
      t.resize(d.nt);
      PMMC_HIP_TRY(ctx, hipMemcpy(t.data(), d.terms + static_cast<size_t>(hs[w].buf) * d.nt,
                                  sizeof(double) * d.nt, hipMemcpyDeviceToHost));
      const int n1 = d.n - 1, n3 = d.n - 2;
      double eb = 0.0, eh = 0.0, ea = 0.0;
      for (int i = 0; i < n1; ++i) eb += t[i];
      for (int i = 0; i < d.np; ++i) eh += t[n1 + i];
      for (int i = 0; i < n3; ++i) ea += t[n1 + d.np + i];
      S.backbone = eb;
      S.pairs = eh;
      S.angular = ea;
      S.total = eb + eh + ea;
      stats[work[w]] = S;
    }
  }
  return PMMC_OK;
}
