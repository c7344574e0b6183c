/*
 * stress.c -- CPU restatement of LooperSolver::stressRefineAnchors' gradient
 * descent (LooperSolver.hpp:5762-5910), in the reference's own form: the
 * gradient scattered into per-bead accumulators loop by loop (springs,
 * pairs, bending windows), energies summed in order in FP64, libm's acosf,
 * float arithmetic with no contraction (built with -ffp-contract=off).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  PINNED BY tests/test_stress.py:
 * bit-exact against the reference's own stressRefineAnchors compiled in
 * place (oracle/_ref/levels_ref, mode "stress"; fixtures in
 * tests/golden/stress_ref.json); the spring and pair gradients against
 * central differences of orc_stress_energy64 and the bending gradient
 * against its forward differences; a two-bead spring relaxing to its target
 * length; the step halving when the energy rises.
 */
#include <math.h>
#include <pthread.h>
#include <string.h>
#include <stdlib.h>

#include "oracle.h"

typedef struct {
  float x, y, z;
} v3;

static v3 v3sub(v3 a, v3 b) {
  v3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
  return r;
}
static float v3len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static v3 ld(const float *x, int i) {
  v3 r = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  return r;
}

/* One evaluation at positions x (LooperSolver.hpp:5771-5859): gradient G
 * [3n] and the three in-order FP64 sums. */
void orc_stress_gradient(int n, const float *x, const float *bb, int np, const int *pi,
                         const int *pj, const float *pt, float ks, float kq, float ka, float *G,
                         double *e_back, double *e_hic, double *e_ang) {
  const float h = 1e-3f;
  memset(G, 0, sizeof(float) * 3 * (size_t)n);
  double eb = 0.0, eh = 0.0, ea = 0.0;
  /* springs */
  for (int i = 0; i + 1 < n; ++i) {
    const float t = bb[i];
    if (t <= 1e-6f) continue;
    const v3 d = v3sub(ld(x, i), ld(x, i + 1));
    float dl = v3len(d);
    if (dl < 1e-9f) dl = 1e-9f;
    const float e = (dl - t) / t;
    const float k = e >= 0.0f ? ks : kq;
    eb += (double)k * e * e;
    const float g = 2.0f * k * (dl - t) / (t * t * dl);
    const float gx = d.x * g, gy = d.y * g, gz = d.z * g;
    G[3 * i] += gx;
    G[3 * i + 1] += gy;
    G[3 * i + 2] += gz;
    G[3 * (i + 1)] -= gx;
    G[3 * (i + 1) + 1] -= gy;
    G[3 * (i + 1) + 2] -= gz;
  }
  /* pairs (weight 1) */
  for (int p = 0; p < np; ++p) {
    const float t = pt[p], w = 1.0f;
    const v3 d = v3sub(ld(x, pi[p]), ld(x, pj[p]));
    float dl = v3len(d);
    if (dl < 1e-9f) dl = 1e-9f;
    const float e = (dl - t) / t;
    eh += (double)w * e * e;
    const float g = 2.0f * w * (dl - t) / (t * t * dl);
    const float gx = d.x * g, gy = d.y * g, gz = d.z * g;
    G[3 * pi[p]] += gx;
    G[3 * pi[p] + 1] += gy;
    G[3 * pi[p] + 2] += gz;
    G[3 * pj[p]] -= gx;
    G[3 * pj[p] + 1] -= gy;
    G[3 * pj[p] + 2] -= gz;
  }
  /* bending: forward differences over the window's beads */
  if (ka > 0.0f) {
    float *X = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    memcpy(X, x, sizeof(float) * 3 * (size_t)n);
    for (int i = 1; i + 1 < n; ++i) {
      const v3 a1 = v3sub(ld(X, i), ld(X, i - 1)), a2 = v3sub(ld(X, i + 1), ld(X, i));
      const float n1 = v3len(a1), n2 = v3len(a2);
      if (n1 < 1e-9f || n2 < 1e-9f) continue;
      float c = (a1.x * a2.x + a1.y * a2.y + a1.z * a2.z) / (n1 * n2);
      if (c > 1.0f) c = 1.0f;
      if (c < -1.0f) c = -1.0f;
      const float th = acosf(c);
      const float E0 = ka * th * th * th;
      ea += (double)E0;
      for (int b = i - 1; b <= i + 1; ++b)
        for (int ax = 0; ax < 3; ++ax) {
          float *xp = &X[3 * b + ax];
          const float saved = *xp;
          *xp = saved + h;
          const v3 w1 = v3sub(ld(X, i), ld(X, i - 1)), w2 = v3sub(ld(X, i + 1), ld(X, i));
          const float m1 = v3len(w1), m2 = v3len(w2);
          *xp = saved;
          if (m1 < 1e-9f || m2 < 1e-9f) continue;
          float cp = (w1.x * w2.x + w1.y * w2.y + w1.z * w2.z) / (m1 * m2);
          if (cp > 1.0f) cp = 1.0f;
          if (cp < -1.0f) cp = -1.0f;
          const float tp = acosf(cp);
          const float Ep = ka * tp * tp * tp;
          G[3 * b + ax] += (Ep - E0) / h;
        }
    }
    free(X);
  }
  *e_back = eb;
  *e_hic = eh;
  *e_ang = ea;
}

/* The descent (LooperSolver.hpp:5762-5910). */
void orc_stress_refine(int n, float *x, const float *bb, int np, const int *pi, const int *pj,
                       const float *pt, float ks, float kq, float ka, int max_iter, float lr,
                       orc_stress_stats *st) {
  memset(st, 0, sizeof(*st));
  st->stop = 3;
  st->lr = lr;
  if (max_iter <= 0 || n < 2) return;
  st->stop = 0;
  float *G = (float *)malloc(sizeof(float) * 3 * (size_t)n);
  double prev = 1e300;
  float cur = lr;
  for (int it = 0; it < max_iter; ++it) {
    double eb, eh, ea;
    orc_stress_gradient(n, x, bb, np, pi, pj, pt, ks, kq, ka, G, &eb, &eh, &ea);
    const double total = eb + eh + ea;
    st->total = total;
    st->backbone = eb;
    st->pairs = eh;
    st->angular = ea;
    if (total > prev && it > 0) {
      cur *= 0.5f;
      if (cur < 1e-6f) {
        st->stop = 2;
        break;
      }
    }
    for (int k = 0; k < 3 * n; ++k) x[k] -= cur * G[k];
    ++st->iterations;
    if (prev - total < 1e-7 * (1.0 < total ? total : 1.0) && it > 20) {
      st->stop = 1;
      break;
    }
    prev = total;
  }
  st->lr = cur;
  free(G);
}

/* FP64 energy at positions x (the same terms, no rounding to float), for the
 * gradient checks. */
double orc_stress_energy64(int n, const float *x, const float *bb, int np, const int *pi,
                           const int *pj, const float *pt, double ks, double kq, double ka) {
  double E = 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    const double t = bb[i];
    if (t <= 1e-6) continue;
    double d = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double u = (double)x[3 * i + k] - x[3 * (i + 1) + k];
      d += u * u;
    }
    d = sqrt(d);
    const double e = (d - t) / t;
    E += (e >= 0.0 ? ks : kq) * e * e;
  }
  for (int p = 0; p < np; ++p) {
    double d = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double u = (double)x[3 * pi[p] + k] - x[3 * pj[p] + k];
      d += u * u;
    }
    d = sqrt(d);
    const double e = (d - pt[p]) / pt[p];
    E += e * e;
  }
  if (ka > 0.0)
    for (int i = 1; i + 1 < n; ++i) {
      double a[3], b[3], na = 0.0, nb = 0.0, dt = 0.0;
      for (int k = 0; k < 3; ++k) {
        a[k] = (double)x[3 * i + k] - x[3 * (i - 1) + k];
        b[k] = (double)x[3 * (i + 1) + k] - x[3 * i + k];
        na += a[k] * a[k];
        nb += b[k] * b[k];
        dt += a[k] * b[k];
      }
      if (na <= 0.0 || nb <= 0.0) continue;
      double c = dt / sqrt(na * nb);
      c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
      const double th = acos(c);
      E += ka * th * th * th;
    }
  return E;
}

/* Checksum of libm's acosf over the floats with bit patterns first ..
 * first + count - 1: sum of splitmix64((bits << 32) | result bits) mod 2^64,
 * on `threads` threads (the device form is pmmc_probe_acosf). */
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

typedef struct {
  uint32_t first, count, t, nt;
  uint64_t sum;
} acos_job;

static void *acos_worker(void *arg) {
  acos_job *j = (acos_job *)arg;
  uint64_t acc = 0;
  for (uint64_t i = j->t; i < j->count; i += j->nt) {
    const uint32_t b = j->first + (uint32_t)i;
    float x, a;
    memcpy(&x, &b, 4);
    a = acosf(x);
    uint32_t r;
    memcpy(&r, &a, 4);
    acc += mix64(((uint64_t)b << 32) | r);
  }
  j->sum = acc;
  return NULL;
}

uint64_t orc_acosf_sum(uint32_t first, uint32_t count, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  acos_job jobs[64];
  for (int t = 0; t < threads; ++t) {
    jobs[t].first = first;
    jobs[t].count = count;
    jobs[t].t = (uint32_t)t;
    jobs[t].nt = (uint32_t)threads;
    pthread_create(&th[t], NULL, acos_worker, &jobs[t]);
  }
  uint64_t s = 0;
  for (int t = 0; t < threads; ++t) {
    pthread_join(th[t], NULL);
    s += jobs[t].sum;
  }
  return s;
}
