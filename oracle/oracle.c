/*
 * oracle.c -- FP64 CPU restatement of pMMC's annealing hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the reference's
 * behaviour; no reference source is included or copied.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* Philox4x32-10, Salmon et al. SC'11, section 3.3 and table 2.        */
/* Multipliers and Weyl key increments are the published constants.   */
/* PINNED BY: the Random123 known-answer vectors (tests/test_oracle.py) */
/* ------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void orc_draw(uint64_t seed, uint32_t move, uint32_t chain, uint32_t segment,
                            uint32_t level, uint32_t sub, uint32_t out[4]) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t ctr[4] = {move, chain, segment, (level << 28) | (sub & 0x0FFFFFFFu)};
  orc_philox4x32_10(ctr, key, out);
}

static inline float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
static inline uint32_t pick(uint32_t x, uint32_t m) {
  return (uint32_t)(((uint64_t)x * (uint64_t)m) >> 32);
}

/* vector3::length() in the reference is a float sqrt of float components
 * (mtxlib); the oracle evaluates the same quantity in double. */
static inline double dist3(const float *a, const float *b) {
  double dx = (double)a[0] - (double)b[0];
  double dy = (double)a[1] - (double)b[1];
  double dz = (double)a[2] - (double)b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

/* ParallelMonteCarloHeatmap.cu:52-61; SphericalContainer.hpp:109-123.
 * PINNED BY: closed form (tests/test_oracle.py::test_sphere_penalty). */
double orc_sphere_penalty(const float *p, double radius) {
  if (radius <= 0.0) return 0.0;
  double d = sqrt((double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2]);
  if (d <= radius) return 0.0;
  return (d - radius) * (d - radius);
}

/* ================================================================== */
/* heatmap level                                                       */
/* ================================================================== */

/* getChromosomeHeatmapBoundary, LooperSolver.hpp:6034-6050: the block
 * [b_i, b_{i+1}-1] that contains p; unchanged (0..n-1) when p lies outside. */
void orc_heatmap_block(const orc_heatmap *P, int p, int *st, int *end) {
  *st = 0;
  *end = P->n - 1;
  if (P->nb <= 0) return;
  if (p < 0 || p > P->bnd[P->nb - 1]) return;
  for (int i = 0; i + 1 < P->nb; ++i) {
    if (P->bnd[i] <= p && p < P->bnd[i + 1]) {
      *st = P->bnd[i];
      *end = P->bnd[i + 1] - 1;
      return;
    }
  }
}

/* term for pair (i = row, m = column) of the local score of bead m:
 * LooperSolver.hpp:3602-3611. valid iff same block, |i-m| >= diag,
 * h[i][m] >= skip_below. */
static inline int hm_valid(const orc_heatmap *P, int i, int m, int st, int end) {
  if (i < st || i > end) return 0;
  if (abs(i - m) < P->diag) return 0;
  if (i == m) return 0;
  if (P->h[(size_t)i * P->n + m] < P->skip_below) return 0;
  return 1;
}

/* PINNED BY: hand cases + brute force full = sum local (test_oracle.py) */
double orc_heatmap_local(const orc_heatmap *P, const float *pos, int moved) {
  int st, end;
  orc_heatmap_block(P, moved, &st, &end);
  double err = 0.0;
  for (int i = st; i <= end; ++i) {
    if (!hm_valid(P, i, moved, st, end)) continue;
    double h = P->h[(size_t)i * P->n + moved];
    double d = dist3(pos + 3 * i, pos + 3 * moved);
    double c = (d - h) / h;
    err += c * c;
  }
  return err;
}

double orc_heatmap_full(const orc_heatmap *P, const float *pos) {
  double err = 0.0;
  for (int m = 0; m < P->n; ++m) err += orc_heatmap_local(P, pos, m);
  return err;
}

/* Exact change of the full score when p moves: terms (i=j, m=p) use h[j][p]
 * and terms (i=p, m=j) use h[p][j]; both blocks are the same partition.
 * PINNED BY: equality with orc_heatmap_full(after) - full(before) on random
 * states (test_oracle.py::test_heatmap_delta_is_full_difference). */
double orc_heatmap_delta(const orc_heatmap *P, const float *pos, int p, const float *newp,
                         double sphere_radius) {
  int st, end;
  orc_heatmap_block(P, p, &st, &end);
  double dE = 0.0;
  for (int j = 0; j < P->n; ++j) {
    if (j == p) continue;
    int v1 = hm_valid(P, j, p, st, end);   /* term of local(p) */
    int jst, jend;
    orc_heatmap_block(P, j, &jst, &jend);
    int v2 = hm_valid(P, p, j, jst, jend); /* term of local(j) */
    if (!v1 && !v2) continue;
    double d0 = dist3(pos + 3 * j, pos + 3 * p);
    double d1 = dist3(pos + 3 * j, newp);
    if (v1) {
      double h = P->h[(size_t)j * P->n + p];
      double a = (d1 - h) / h, b = (d0 - h) / h;
      dE += a * a - b * b;
    }
    if (v2) {
      double h = P->h[(size_t)p * P->n + j];
      double a = (d1 - h) / h, b = (d0 - h) / h;
      dE += a * a - b * b;
    }
  }
  dE += orc_sphere_penalty(newp, sphere_radius) - orc_sphere_penalty(pos + 3 * p, sphere_radius);
  return dE;
}

static double heatmap_total(const orc_heatmap *P, const float *pos, double r) {
  double e = orc_heatmap_full(P, pos);
  if (r > 0.0)
    for (int i = 0; i < P->n; ++i) e += orc_sphere_penalty(pos + 3 * i, r);
  return e;
}

/* ================================================================== */
/* arc level                                                           */
/* ================================================================== */

static inline double arc_term(double d, double e, double ks, double kq) {
  double diff = (d - e) / e;
  return diff * diff * (diff >= 0.0 ? ks : kq);
}

/* repulsion of pairs with no arc (e < 0): LooperSolver.hpp:3228-3231,
 * ParallelMonteCarloArcs.cu:483-486 (1/max(d,1e-6)). */
double orc_arcs_repulsion(const orc_arcs *P, const float *pos) {
  double sc = 0.0;
  int n = P->n;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (P->ed[(size_t)i * n + j] < 0.0f) {
        double d = dist3(pos + 3 * i, pos + 3 * j);
        sc += 1.0 / (d > 1e-6 ? d : 1e-6);
      }
    }
  return sc;
}

static double arcs_pairs_and_loops(const orc_arcs *P, const float *pos) {
  double sc = 0.0;
  int n = P->n;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double e = P->ed[(size_t)i * n + j];
      if (e < 1e-6) continue;
      sc += arc_term(dist3(pos + 3 * i, pos + 3 * j), e, P->ks, P->kq);
    }
  for (int k = 0; k < P->nloops; ++k) {
    int a = P->lpairs[2 * k], b = P->lpairs[2 * k + 1];
    double d = dist3(pos + 3 * a, pos + 3 * b) - P->lparams[2 * k + 1];
    sc += P->lparams[2 * k] * d * d;
  }
  return sc;
}

/* PINNED BY: hand cases (stretch/squeeze branch, repulsion, skip rules) and
 * brute force on n<=8 (test_oracle.py) */
double orc_arcs_total(const orc_arcs *P, const float *pos, double sphere_radius) {
  double sc = orc_arcs_repulsion(P, pos) + arcs_pairs_and_loops(P, pos);
  if (sphere_radius > 0.0)
    for (int i = 0; i < P->n; ++i) sc += orc_sphere_penalty(pos + 3 * i, sphere_radius);
  return sc;
}

/* Local score of bead p: dense column scan as LooperSolver.hpp:3250-3281
 * (e < 1e-6 skipped, no repulsion), plus loops touching p (:3355-3371) and
 * sphere(p).  Row semantics of the GPU kernel (.cu:294-320) coincide for
 * symmetric ed.  PINNED BY: local(after)-local(before) == total difference
 * minus repulsion change (test_oracle.py). */
double orc_arcs_local(const orc_arcs *P, const float *pos, int p, const float *pp,
                      double sphere_radius) {
  const float *x = pp ? pp : pos + 3 * p;
  int n = P->n;
  double sc = 0.0;
  for (int i = 0; i < n; ++i) {
    if (i == p) continue;
    double e = P->ed[(size_t)i * n + p];
    if (e < 1e-6) continue;
    sc += arc_term(dist3(pos + 3 * i, x), e, P->ks, P->kq);
  }
  for (int k = 0; k < P->nloops; ++k) {
    int a = P->lpairs[2 * k], b = P->lpairs[2 * k + 1];
    if (a != p && b != p) continue;
    const float *xa = (a == p) ? x : pos + 3 * a;
    const float *xb = (b == p) ? x : pos + 3 * b;
    double d = dist3(xa, xb) - P->lparams[2 * k + 1];
    sc += P->lparams[2 * k] * d * d;
  }
  sc += orc_sphere_penalty(x, sphere_radius);
  return sc;
}

/* ================================================================== */
/* smooth level                                                        */
/* ================================================================== */

static inline double bond_term(const float *a, const float *b, float dtn, double ks, double kq) {
  double d0 = dtn;
  if (d0 < 1e-6) d0 = 1e-6;
  double diff = (dist3(a, b) - d0) / d0;
  return diff * diff * (diff >= 0.0 ? ks : kq);
}

/* angle term between bond vectors u = a-b and v = b-c (vertex b):
 * ((1 - cos)/2)^3 * ka, skipped when either length <= 1e-6, cos clamped;
 * ParallelMonteCarloSmooth.cu:126-147 and common.cpp:39-46. */
static inline double angle_term(const float *a, const float *b, const float *c, double ka) {
  double ux = (double)a[0] - b[0], uy = (double)a[1] - b[1], uz = (double)a[2] - b[2];
  double vx = (double)b[0] - c[0], vy = (double)b[1] - c[1], vz = (double)b[2] - c[2];
  double lu = sqrt(ux * ux + uy * uy + uz * uz), lv = sqrt(vx * vx + vy * vy + vz * vz);
  if (lu <= 1e-6 || lv <= 1e-6) return 0.0;
  double cs = (ux * vx + uy * vy + uz * vz) / (lu * lv);
  if (cs < -1.0) cs = -1.0;
  if (cs > 1.0) cs = 1.0;
  double ang = (1.0 - cs) * 0.5;
  return ang * ang * ang * ka;
}

/* PINNED BY: straight chain -> 0 angle energy; right angle -> ka/8; bonds at
 * rest -> 0 (test_oracle.py) */
double orc_smooth_full(const orc_smooth *P, const float *pos, double sphere_radius, int weighted) {
  int n = P->n;
  double sca = 0.0, scb = 0.0, sph = 0.0;
  for (int i = 0; i + 1 < n; ++i)
    sca += bond_term(pos + 3 * i, pos + 3 * (i + 1), P->dtn[i], P->ks, P->kq);
  for (int i = 1; i + 1 < n; ++i)
    scb += angle_term(pos + 3 * (i - 1), pos + 3 * i, pos + 3 * (i + 1), P->ka);
  if (sphere_radius > 0.0)
    for (int i = 0; i < n; ++i) sph += orc_sphere_penalty(pos + 3 * i, sphere_radius);
  if (weighted) return sca * P->wd + scb * P->wa + sph;
  return sca + scb + sph;
}

/* bonds (p-1,p), (p,p+1) and angles at p-1, p, p+1, unweighted, + sphere(p).
 * PINNED BY: local(after)-local(before) == unweighted full difference on
 * random chains (test_oracle.py) */
double orc_smooth_local(const orc_smooth *P, const float *pos, int p, const float *pp,
                        double sphere_radius) {
  int n = P->n;
  const float *x = pp ? pp : pos + 3 * p;
#define SP(i) ((i) == p ? x : pos + 3 * (i))
  double sc = 0.0;
  if (p > 0) sc += bond_term(SP(p - 1), x, P->dtn[p - 1], P->ks, P->kq);
  if (p < n - 1) sc += bond_term(x, SP(p + 1), P->dtn[p], P->ks, P->kq);
  for (int v = p - 1; v <= p + 1; ++v) {
    if (v < 1 || v + 1 >= n) continue;
    sc += angle_term(SP(v - 1), SP(v), SP(v + 1), P->ka);
  }
#undef SP
  sc += orc_sphere_penalty(x, sphere_radius);
  return sc;
}

/* ================================================================== */
/* generic chain                                                       */
/* ================================================================== */

typedef struct {
  int level;
  const void *prob;
  int n;
  const uint8_t *fixed;
  double r;               /* sphere radius */
  double offset;          /* smooth: weighted(x0) - unweighted(x0); arcs: repulsion(x0) */
  int full_rescore;
} chain_ctx;

static double ctx_total(const chain_ctx *c, const float *pos) {
  switch (c->level) {
    case ORC_LEVEL_HEATMAP: return heatmap_total((const orc_heatmap *)c->prob, pos, c->r);
    case ORC_LEVEL_ARCS: {
      const orc_arcs *P = (const orc_arcs *)c->prob;
      double e = c->offset + arcs_pairs_and_loops(P, pos);
      if (c->r > 0.0)
        for (int i = 0; i < P->n; ++i) e += orc_sphere_penalty(pos + 3 * i, c->r);
      return e;
    }
    default:
      return c->offset + orc_smooth_full((const orc_smooth *)c->prob, pos, c->r, 0);
  }
}

static double ctx_delta(const chain_ctx *c, float *pos, int p, const float *np) {
  switch (c->level) {
    case ORC_LEVEL_HEATMAP: {
      const orc_heatmap *P = (const orc_heatmap *)c->prob;
      if (c->full_rescore) {
        /* the reference's CPU cost structure: rescore everything */
        float save[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        double before = heatmap_total(P, pos, c->r);
        pos[3 * p] = np[0]; pos[3 * p + 1] = np[1]; pos[3 * p + 2] = np[2];
        double after = heatmap_total(P, pos, c->r);
        pos[3 * p] = save[0]; pos[3 * p + 1] = save[1]; pos[3 * p + 2] = save[2];
        return after - before;
      }
      return orc_heatmap_delta(P, pos, p, np, c->r);
    }
    case ORC_LEVEL_ARCS: {
      const orc_arcs *P = (const orc_arcs *)c->prob;
      return orc_arcs_local(P, pos, p, np, c->r) - orc_arcs_local(P, pos, p, NULL, c->r);
    }
    default: {
      const orc_smooth *P = (const orc_smooth *)c->prob;
      return orc_smooth_local(P, pos, p, np, c->r) - orc_smooth_local(P, pos, p, NULL, c->r);
    }
  }
}

/* |local before| + |local after| of moving bead p to np: the magnitude an
 * FP32 evaluation of the move's ΔE is formed from */
static double ctx_scale(const chain_ctx *c, float *pos, int p, const float *np) {
  switch (c->level) {
    case ORC_LEVEL_HEATMAP: {
      const orc_heatmap *P = (const orc_heatmap *)c->prob;
      const double a = orc_heatmap_local(P, pos, p);
      float save[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
      pos[3 * p] = np[0]; pos[3 * p + 1] = np[1]; pos[3 * p + 2] = np[2];
      const double b = orc_heatmap_local(P, pos, p);
      pos[3 * p] = save[0]; pos[3 * p + 1] = save[1]; pos[3 * p + 2] = save[2];
      return 2.0 * (fabs(a) + fabs(b));
    }
    case ORC_LEVEL_ARCS: {
      const orc_arcs *P = (const orc_arcs *)c->prob;
      return fabs(orc_arcs_local(P, pos, p, NULL, c->r)) + fabs(orc_arcs_local(P, pos, p, np, c->r));
    }
    default: {
      const orc_smooth *P = (const orc_smooth *)c->prob;
      return fabs(orc_smooth_local(P, pos, p, NULL, c->r)) + fabs(orc_smooth_local(P, pos, p, np, c->r));
    }
  }
}

int64_t orc_anneal_chain(int level, const void *problem, const orc_schedule *s,
                         uint64_t seed, uint32_t segment, uint32_t chain, int pick_shift,
                         float *pos, orc_stats *st, uint8_t *accept_trace,
                         int full_rescore, double tie_rel) {
  chain_ctx c;
  memset(&c, 0, sizeof(c));
  c.level = level;
  c.prob = problem;
  c.r = s->sphere_radius;
  c.full_rescore = full_rescore;
  if (level == ORC_LEVEL_HEATMAP) {
    c.n = ((const orc_heatmap *)problem)->n;
    c.fixed = ((const orc_heatmap *)problem)->fixed;
  } else if (level == ORC_LEVEL_ARCS) {
    const orc_arcs *P = (const orc_arcs *)problem;
    c.n = P->n;
    c.fixed = P->fixed;
    c.offset = orc_arcs_repulsion(P, pos);
  } else {
    const orc_smooth *P = (const orc_smooth *)problem;
    c.n = P->n;
    c.fixed = P->fixed;
    c.offset = orc_smooth_full(P, pos, c.r, 1) - orc_smooth_full(P, pos, c.r, 0);
  }

  memset(st, 0, sizeof(*st));
  int n = c.n;
  int *freel = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int nfree = 0;
  for (int i = 0; i < n; ++i)
    if (!c.fixed || !c.fixed[i]) freel[nfree++] = i;

  double E = ctx_total(&c, pos);
  double T = s->T0;
  float step = s->step;
  double ms = E;
  int64_t iters = 0, acc = 0;
  int32_t msucc = 0, misses = 0, milestones = 0;
  double best = E;
  int reason = 0;
  int32_t ties = 0;
  int64_t first_tie = -1;
  double drift = 0.0;  /* tie model: tolerance summed over accepted moves since a re-sum */

  if (!(E == E) || E < s->guard_energy) reason = 4;
  else if (nfree == 0) reason = 5;

  while (reason == 0) {
    if (iters >= s->max_moves) { reason = 3; break; }
    for (int k = 0; k < s->block; ++k) {
      uint32_t r0[4], rb[4];
      uint32_t m = (uint32_t)iters;
      orc_draw(seed, m >> 2, chain >> pick_shift, segment, (uint32_t)level, 1, rb);
      int p = freel[pick(rb[m & 3], (uint32_t)nfree)];
      orc_draw(seed, m, chain, segment, (uint32_t)level, 0, r0);
      float d[3];
      d[0] = (2.0f * u01(r0[0]) - 1.0f) * step;
      d[1] = (2.0f * u01(r0[1]) - 1.0f) * step;
      d[2] = s->use2d ? 0.0f : (2.0f * u01(r0[2]) - 1.0f) * step;
      float np[3] = {pos[3 * p] + d[0], pos[3 * p + 1] + d[1], pos[3 * p + 2] + d[2]};
      double dE = ctx_delta(&c, pos, p, np);
      double En = E + dE;
      int ok = s->strict_accept ? (En < E) : (En <= E);
      double tol = 0.0;
      int tie = 0;
      if (tie_rel > 0.0) {
        tol = tie_rel * ctx_scale(&c, pos, p, np) + ORC_TIE_ABS;
        tie = fabs(dE) <= tol;
      }
      if (!ok && T > 0.0 && (!s->positive_guard || E > 0.0)) {
        const double arg = s->tjc * (En / E) / T;
        double tp = s->tjs * exp(-arg);
        const double U = (double)u01(r0[3]);
        ok = U < tp;
        if (tie_rel > 0.0) {
          const double B = tie_rel * (1.0 + fabs(arg)) + (s->tjc / T) * (tol + drift) / fabs(E);
          tie = tie || fabs(U - tp) <= B * tp;
        }
      }
      if (tie) {
        if (first_tie < 0) first_tie = iters;
        ++ties;
      }
      if (ok && tie_rel > 0.0) drift += tol;
      if (ok) {
        pos[3 * p] = np[0]; pos[3 * p + 1] = np[1]; pos[3 * p + 2] = np[2];
        E = En;
        ++msucc;
        ++acc;
        if (E < best) best = E;
      }
      if (accept_trace) accept_trace[iters] = (uint8_t)ok;
      if (!s->cool_per_block) T *= s->dt;
      ++iters;
    }
    if (s->cool_per_block) T *= s->dt;
    step *= s->step_decay;
    if ((iters % s->milestone_period) < s->block) {
      ++milestones;
      if (s->resum > 0 && milestones % s->resum == 0) {
        E = ctx_total(&c, pos);
        drift = 0.0;
      }
      if (!isfinite(E)) { reason = 6; break; }
      int no_imp = (E > s->improvement * ms && msucc < s->min_successes) || E < s->small_energy ||
                   (s->ratio_stop > 0.0 && ms > 1e-10 && E / ms > s->ratio_stop);
      if (s->stop_rule == ORC_STOP_FIRST_FAIL) {
        if (no_imp) { reason = E < s->small_energy ? 2 : 1; break; }
      } else {
        if (no_imp) ++misses;
        else if (s->stop_rule == ORC_STOP_MISSES_RESET) misses = 0;
        if (E < s->small_energy) { reason = 2; break; }
        if (misses >= s->fails_threshold) { reason = 1; break; }
      }
      ms = E;
      msucc = 0;
    }
  }
  free(freel);
  st->energy = E;
  st->best_energy = best;
  st->T = T;
  st->moves = iters;
  st->accepted = acc;
  st->milestones = milestones;
  st->stop_reason = reason;
  st->near_ties = ties;
  st->first_near_tie = first_tie;
  return iters;
}

/* ------------------------------------------------------------------ */
/* many chains on host threads (CPU baseline)                          */
/* ------------------------------------------------------------------ */
typedef struct {
  int level;
  const void *problem;
  const orc_schedule *s;
  uint64_t seed;
  uint32_t segment, chain0;
  int n_chains, n, full_rescore, stride, pick_shift;
  double tie_rel;
  int tid;
  const float *pos0;
  float *pos_out;
  orc_stats *stats;
} many_arg;

static void *many_worker(void *vp) {
  many_arg *a = (many_arg *)vp;
  for (int c = a->tid; c < a->n_chains; c += a->stride) {
    float *p = a->pos_out + (size_t)c * a->n * 3;
    memcpy(p, a->pos0, sizeof(float) * 3 * (size_t)a->n);
    orc_anneal_chain(a->level, a->problem, a->s, a->seed, a->segment, a->chain0 + (uint32_t)c,
                     a->pick_shift, p, &a->stats[c], NULL, a->full_rescore, a->tie_rel);
  }
  return NULL;
}

double orc_anneal_many(int level, const void *problem, const orc_schedule *s,
                       uint64_t seed, uint32_t segment, uint32_t chain0, int n_chains,
                       int pick_shift, const float *pos0, float *pos_out, orc_stats *stats,
                       int n_threads, int full_rescore, double tie_rel) {
  int n = level == ORC_LEVEL_HEATMAP ? ((const orc_heatmap *)problem)->n
        : level == ORC_LEVEL_ARCS    ? ((const orc_arcs *)problem)->n
                                     : ((const orc_smooth *)problem)->n;
  if (n_threads < 1) n_threads = 1;
  if (n_threads > n_chains) n_threads = n_chains;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  many_arg *args = (many_arg *)malloc(sizeof(many_arg) * n_threads);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < n_threads; ++t) {
    many_arg a = {level, problem, s, seed, segment, chain0, n_chains, n, full_rescore,
                  n_threads, pick_shift, tie_rel, t, pos0, pos_out, stats};
    args[t] = a;
    pthread_create(&th[t], NULL, many_worker, &args[t]);
  }
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  free(args);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ================================================================== */
/* heatmap sweep: the reference GPU heatmap kernel's algorithm          */
/* (ParallelMonteCarloHeatmap.cu:217-363) with races replaced by a     */
/* synchronous (Jacobi) sweep over a snapshot of the positions.        */
/* PINNED BY (tests/test_oracle_sweep.py): an independent Python       */
/* enumeration of the sweep's definition on tiny cases (n <= 4,        */
/* proposers <= 2, inner <= 4, several iterations, fixed beads,        */
/* sphere); the invariants the definition fixes (a bead moves only to  */
/* a point of strictly lower snapshot-local score, fixed beads never   */
/* move, at T0 = 0 no proposer accepts an uphill move); the argmin's   */
/* tie rule (orc_sweep_argmin, hand cases); the stop rule's truth      */
/* table over the energy trace (misses never reset).  Parity against   */
/* the reference GPU kernel itself stays unpinned (CUDA, racy).        */
/* ================================================================== */

/* lowest score, ties to the lowest index (proposer index = replica * 64 +
 * lane, so this is "lowest lane, then lowest replica") */
int orc_sweep_argmin(const double *sc, int n) {
  int b = 0;
  for (int r = 1; r < n; ++r)
    if (sc[r] < sc[b]) b = r;
  return b;
}

/* local score with the GPU skip threshold, bead b at x, against snapshot */
static double sweep_local(const orc_heatmap *P, const float *snap, int b, const float *x) {
  int st, end;
  orc_heatmap_block(P, b, &st, &end);
  double err = 0.0;
  for (int i = st; i <= end; ++i) {
    if (!hm_valid(P, i, b, st, end)) continue;
    double h = P->h[(size_t)i * P->n + b];
    double c = dist3(snap + 3 * i, x) / h - 1.0;
    err += c * c;
  }
  return err;
}

/* near-tie bound of a comparison between two FP32-evaluated scores */
static inline int near(double a, double b, double rel) {
  return fabs(a - b) <= rel * (fabs(a) + fabs(b)) + ORC_TIE_ABS;
}

int orc_heatmap_sweep(const orc_heatmap *P, const orc_sweep_schedule *s, uint64_t seed,
                      uint32_t segment, float *pos, orc_stats *st, double tie_rel,
                      uint8_t *tie_beads, double *e_trace) {
  int n = P->n;
  memset(st, 0, sizeof(*st));
  st->first_near_tie = -1;
  if (tie_beads) memset(tie_beads, 0, (size_t)n);
  float *snap = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
  double *sprev = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *psc = (double *)malloc(sizeof(double) * (size_t)(s->proposers > 0 ? s->proposers : 1));
  float *px = (float *)malloc(sizeof(float) * 3 * (size_t)(s->proposers > 0 ? s->proposers : 1));
  uint8_t *itie = (uint8_t *)malloc((size_t)(n > 0 ? n : 1));
  double E = orc_heatmap_full(P, pos);
  if (e_trace) e_trace[0] = E;
  for (int b = 0; b < n; ++b) sprev[b] = E; /* first comparison against the total (Q3) */
  double T = s->T0, ms = E;
  float step = s->step;
  int misses = 0, reason = 0, it = 0;
  int64_t acc = 0, moves = 0;
  if (!(E == E) || E < 1e-6) reason = 4;
  while (reason == 0) {
    if (it >= s->max_iters) { reason = 3; break; }
    memcpy(snap, pos, sizeof(float) * 3 * (size_t)n);
    memset(itie, 0, (size_t)n);
    int any_tie = 0;
    for (int b = 0; b < n; ++b) {
      if (P->fixed && P->fixed[b]) continue;
      for (int r = 0; r < s->proposers; ++r) {
        float x[3] = {snap[3 * b], snap[3 * b + 1], snap[3 * b + 2]};
        double sp = sprev[b];
        for (int k = 0; k < s->inner; ++k) {
          uint32_t r0[4];
          uint32_t mv = (uint32_t)it * (uint32_t)s->inner + (uint32_t)k;
          uint32_t ch = (uint32_t)b * (uint32_t)s->proposers + (uint32_t)r;
          orc_draw(seed, mv, ch, segment, ORC_LEVEL_SWEEP, 0, r0);
          float d[3];
          d[0] = (2.0f * u01(r0[0]) - 1.0f) * step;
          d[1] = (2.0f * u01(r0[1]) - 1.0f) * step;
          d[2] = s->use2d ? 0.0f : (2.0f * u01(r0[2]) - 1.0f) * step;
          float y[3] = {x[0] + d[0], x[1] + d[1], x[2] + d[2]};
          if (s->use2d) y[2] = x[2];
          double sn = sweep_local(P, snap, b, y) + orc_sphere_penalty(y, s->sphere_radius);
          int ok = sn <= sp;
          int tie = tie_rel > 0.0 && near(sn, sp, tie_rel);
          if (!ok && T > 0.0) {
            const double arg = s->tjc * (sn / sp) / T;
            double tp = s->tjs * exp(-arg);
            const double U = (double)u01(r0[3]);
            ok = U < tp;
            /* the FP32 ratio carries both scores' errors */
            if (tie_rel > 0.0) tie = tie || fabs(U - tp) <= tie_rel * (1.0 + 2.0 * fabs(arg)) * tp;
          }
          if (tie) itie[b] = 1;
          ++moves;
          if (ok) {
            x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
            sp = sn;
            ++acc;
          }
        }
        psc[r] = sp;
        px[3 * r] = x[0]; px[3 * r + 1] = x[1]; px[3 * r + 2] = x[2];
      }
      const int br = orc_sweep_argmin(psc, s->proposers);
      if (tie_rel > 0.0)
        for (int r = 0; r < s->proposers; ++r)
          if (r != br && near(psc[r], psc[br], tie_rel)) itie[b] = 1;
      const float *best_x = px + 3 * br;
      double lc = sweep_local(P, snap, b, best_x);
      double lg = sweep_local(P, snap, b, snap + 3 * b);
      if (tie_rel > 0.0 && fabs(lc - lg) <= 1e-12 * (lc + lg)) itie[b] = 1;
      if (lc < lg) { pos[3 * b] = best_x[0]; pos[3 * b + 1] = best_x[1]; pos[3 * b + 2] = best_x[2]; }
      sprev[b] = psc[br];
      any_tie |= itie[b];
    }
    T *= s->dt;
    step *= s->step_decay;
    ++it;
    E = orc_heatmap_full(P, pos);
    if (e_trace) e_trace[it] = E;
    /* the stop test compares FP64 sums formed in another order */
    const int stop_tie = tie_rel > 0.0 &&
                         (fabs(E - s->improvement * ms) <= 1e-12 * fabs(E) ||
                          fabs(E - s->small_energy) <= 1e-12 * fabs(E));
    if (any_tie || stop_tie) {
      ++st->near_ties;
      if (st->first_near_tie < 0) {
        st->first_near_tie = it - 1;
        if (tie_beads) {
          if (stop_tie) memset(tie_beads, 1, (size_t)n);
          else memcpy(tie_beads, itie, (size_t)n);
        }
      }
    }
    if (!isfinite(E)) { reason = 6; break; }
    if (E > s->improvement * ms) ++misses;   /* never reset (Q4) */
    if (E < s->small_energy) { reason = 2; break; }
    if (misses >= s->fails_threshold) { reason = 1; break; }
    ms = E;
  }
  free(snap);
  free(sprev);
  free(psc);
  free(px);
  free(itie);
  st->energy = E;
  st->best_energy = E;
  st->T = T;
  st->moves = moves;
  st->accepted = acc;
  st->milestones = it;
  st->stop_reason = reason;
  return it;
}

/* ================================================================== */
/* arc marking: ParallelMarkArcs.cu:26-64 and InteractionArcs.hpp:226-235 */
/* Anchor::length (Anchor.hpp:81-85): 0 when start == end == 0, else      */
/* end - start + 1.  PINNED BY: closed form on hand cases (overlaps,     */
/* zero-length, boundaries) in test_oracle.py.                            */
/* ================================================================== */
void orc_mark_arcs(const int32_t *anc_s, const int32_t *anc_e, int32_t A,
                   const int32_t *arc_s, const int32_t *arc_e, int32_t R,
                   int32_t *out_s, int32_t *out_e) {
  for (int i = 0; i < R; ++i) {
    int st = -1, en = -1;
    for (int j = 0; j < A; ++j) {
      int len = (anc_s[j] == 0 && anc_e[j] == 0) ? 0 : anc_e[j] - anc_s[j] + 1;
      if (len <= 1) continue;
      if (arc_s[i] >= anc_s[j] && arc_s[i] <= anc_e[j]) st = j;
      if (arc_e[i] >= anc_s[j] && arc_e[i] <= anc_e[j]) en = j;
    }
    out_s[i] = st;
    out_e[i] = en;
  }
}
