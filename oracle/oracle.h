/*
 * oracle.h -- FP64 CPU restatement of pMMC's simulated-annealing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (pmmc_amd/) links,
 * loads or calls this library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / the CPU
 * baseline being timed.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's src/ directory).  Pinning status per function is listed in
 * the "PINNED BY" notes; a function without a pin says "parity unpinned".
 *
 * Conventions shared with the HIP kernels (the stream specification):
 *   - positions: float xyz interleaved, pos[3*i+k]; stored and updated in FP32
 *     exactly as the reference stores vector3/float3.
 *   - energies: accumulated in FP64 here (the kernels use FP32 for per-move
 *     deltas and FP64 for the running total, re-summation and temperature).
 *   - random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11
 *     "Parallel random numbers: as easy as 1, 2, 3").  Stream layout:
 *       key     = { seed & 0xffffffff, seed >> 32 }
 *       counter = { move, chain, segment, (level << 28) | sub }
 *     per chain, move m:   Philox({m, chain, segment, level<<28 | 0}) = {dx, dy, dz, U}
 *     bead of move m:      Philox({m >> 2, chain >> pick_shift, segment,
 *                                  level<<28 | 1})[m & 3]
 *     (pick_shift = 6: the 64 chains of one wavefront share bead picks.)
 *     u01(x) = (x >> 8) * 2^-24 ; pick(x, m) = (x * m) >> 32 (64-bit product).
 *     displacement component = (2*u01 - 1) * step   (FP32, as the reference).
 */
#ifndef PMMC_ORACLE_H
#define PMMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- levels (Philox level field) ---- */
enum { ORC_LEVEL_HEATMAP = 1, ORC_LEVEL_ARCS = 2, ORC_LEVEL_SMOOTH = 3,
       ORC_LEVEL_SWEEP = 4, ORC_LEVEL_INIT = 5 };

/* ---- stop rules (see orc_schedule) ---- */
enum { ORC_STOP_FIRST_FAIL = 0,   /* CPU annealers: break at first failing milestone */
       ORC_STOP_MISSES_RESET = 1, /* GPU arcs: misses reset on improvement          */
       ORC_STOP_MISSES_ACCUM = 2  /* GPU heatmap / smooth: misses never reset       */ };

/* Schedule of one Metropolis chain.  Field-for-field the same meaning as the
 * product's pmmc_schedule.  The presets that reproduce each reference
 * annealer are built by the tests (tests/schedules.py). */
typedef struct {
  double T0;             /* initial temperature                                  */
  double dt;             /* cooling factor                                       */
  double improvement;    /* milestone improvement threshold                      */
  double small_energy;   /* stop when E < small_energy                           */
  double ratio_stop;     /* >0: E/ms > ratio_stop counts as no improvement       */
  double guard_energy;   /* skip annealing when initial E < guard (or NaN)       */
  float tjs, tjc;        /* tempJumpScale, tempJumpCoef                          */
  float step;            /* displacement half width (already scaled by caller)   */
  float step_decay;      /* step *= step_decay after every block                 */
  float sphere_radius;   /* 0 = off                                              */
  int32_t block;         /* moves per inner block                                */
  int32_t milestone_period; /* milestone when (iters % period) < block         */
  int32_t min_successes; /* successes threshold in the no-improvement test       */
  int32_t fails_threshold;  /* misses threshold (MISSES_* rules)               */
  int32_t stop_rule;     /* ORC_STOP_*                                            */
  int32_t strict_accept; /* 1: downhill test is E_new <  E ; 0: E_new <= E       */
  int32_t positive_guard;/* 1: uphill acceptance requires E > 0                  */
  int32_t cool_per_block;/* 1: T *= dt once per block; 0: once per move          */
  int32_t use2d;         /* 1: dz = 0                                            */
  int32_t resum;         /* K > 0: re-sum total energy at every K-th milestone   */
  int64_t max_moves;     /* cap, checked before each block                       */
} orc_schedule;

typedef struct {
  double energy;         /* final running total (FP64)          */
  double best_energy;    /* lowest running total seen           */
  double T;              /* final temperature                   */
  int64_t moves;         /* counted moves                       */
  int64_t accepted;
  int32_t milestones;
  int32_t stop_reason;   /* 0 running,1 stop rule,2 small,3 cap,4 guard,5 no free bead,6 non-finite */
  int32_t near_ties;     /* decisions within FP32 error of their threshold (track_ties) */
  int32_t pad_;
  int64_t first_near_tie;  /* move index of the first, -1 if none          */
} orc_stats;

/* ---- Philox4x32-10 ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- sphere penalty: ParallelMonteCarloHeatmap.cu:52-61, SphericalContainer.hpp:109-123 */
double orc_sphere_penalty(const float *p, double radius);

/* ================= heatmap level ================= */
/* Heatmap problem (all host arrays).  h is heatmap_dist, n x n row-major. */
typedef struct {
  int32_t n;
  const float *h;          /* [n*n]                      */
  int32_t diag;            /* heatmap diagonal_size      */
  const int32_t *bnd;      /* chromosome boundaries      */
  int32_t nb;              /* number of boundaries       */
  float skip_below;        /* 1e-6 (CPU) / 1e-3 (GPU)   */
  const uint8_t *fixed;    /* [n] or NULL                */
} orc_heatmap;

/* local(m): LooperSolver.hpp:3583-3617 (moved >= 0 branch) */
double orc_heatmap_local(const orc_heatmap *P, const float *pos, int moved);
/* full: LooperSolver.hpp:3592-3594 (sum of local over all beads) */
double orc_heatmap_full(const orc_heatmap *P, const float *pos);
/* exact change of full when bead p moves to newp */
double orc_heatmap_delta(const orc_heatmap *P, const float *pos, int p, const float *newp,
                         double sphere_radius);
/* getChromosomeHeatmapBoundary: LooperSolver.hpp:6034-6050 */
void orc_heatmap_block(const orc_heatmap *P, int p, int *st, int *end);

/* ================= arc level ================= */
typedef struct {
  int32_t n;
  const float *ed;         /* expected distances [n*n], -1 = no arc      */
  int32_t nloops;
  const int32_t *lpairs;   /* [2*nloops] active-region indices           */
  const float *lparams;    /* [2*nloops] {stiffness, eq_distance}         */
  float ks, kq;            /* springConstantStretchArcs / SqueezeArcs     */
  const uint8_t *fixed;    /* [n] or NULL                                 */
} orc_arcs;

/* total with repulsion: LooperSolver.hpp:3215-3246 + loops :3339-3353 */
double orc_arcs_total(const orc_arcs *P, const float *pos, double sphere_radius);
/* repulsion part only (the part that the running total freezes) */
double orc_arcs_repulsion(const orc_arcs *P, const float *pos);
/* local of bead p at position pp (NULL = current): :3250-3281 + :3355-3371 */
double orc_arcs_local(const orc_arcs *P, const float *pos, int p, const float *pp,
                      double sphere_radius);

/* ================= smooth level ================= */
typedef struct {
  int32_t n;
  const float *dtn;        /* dist_to_next [n]            */
  float ks, kq, ka;        /* stretch, squeeze, angular   */
  float wd, wa;            /* weightDistSmooth, weightAngleSmooth */
  const uint8_t *fixed;    /* [n] or NULL                 */
} orc_smooth;

/* weighted full energy: LooperSolver.hpp:3414-3451, ParallelMonteCarloSmooth.cu:157-209 */
double orc_smooth_full(const orc_smooth *P, const float *pos, double sphere_radius, int weighted);
/* unweighted local window of bead p at position pp: :3453-3495, .cu:86-154 */
double orc_smooth_local(const orc_smooth *P, const float *pos, int p, const float *pp,
                        double sphere_radius);

/* ================= annealer ================= */
/* One Metropolis chain following orc_schedule; pos is updated in place.
 * heatmap: ΔE = exact change of the full score (+ sphere);
 *   full_rescore=1 recomputes the full O(N^2) score every move, which is what
 *   LooperSolver::MonteCarloHeatmap (LooperSolver.hpp:1488) does -- used to
 *   time the reference's CPU cost structure, same result up to rounding.
 * arcs: running total = repulsion(initial) + arcs + loops + sphere
 *   (MonteCarloArcs LooperSolver.hpp:4670-4796; MonteCarloArcsKernel
 *   ParallelMonteCarloArcs.cu:407-675).
 * smooth: running total = weighted(initial) + Σ unweighted local deltas
 *   (MonteCarloArcsSmooth LooperSolver.hpp:4798-5047; .cu:231-349).
 * accept_trace (nullable, length >= moves): 1 per accepted move.
 * tie_rel > 0 counts near ties (st->near_ties, st->first_near_tie):
 *   decisions that an FP32 evaluation of the move with relative error
 *   tie_rel could take the other way.  A move is one when |dE| <= tie_rel *
 *   (|local before| + |local after|) + 1e-7, or, when the Metropolis test
 *   decides, when |U - p| <= B * p with p the acceptance probability and B
 *   its relative error carried through arg = tjc (E'/E) / T: tie_rel
 *   (1 + |arg|) plus (tjc / T) times the ΔE error and the running total's
 *   drift (the ΔE errors summed over accepted moves since the last
 *   re-summation), over |E|.
 * Returns the number of counted moves. */
#define ORC_TIE_ABS 1e-7
int64_t orc_anneal_chain(int level, const void *problem, const orc_schedule *s,
                         uint64_t seed, uint32_t segment, uint32_t chain, int pick_shift,
                         float *pos, orc_stats *st, uint8_t *accept_trace,
                         int full_rescore, double tie_rel);

/* Run n_chains chains (chain ids chain0..chain0+n_chains-1), each from the
 * same initial positions, on n_threads host threads.  pos_out [n_chains*n*3]
 * and stats [n_chains] are written.  Returns wall seconds spent annealing. */
double orc_anneal_many(int level, const void *problem, const orc_schedule *s,
                       uint64_t seed, uint32_t segment, uint32_t chain0, int n_chains,
                       int pick_shift,
                       const float *pos0, float *pos_out, orc_stats *stats,
                       int n_threads, int full_rescore, double tie_rel);

/* ================= heatmap sweep (reference GPU heatmap kernel) ================= */
/* MonteCarloHeatmapKernel semantics (ParallelMonteCarloHeatmap.cu:217-363) with
 * the inter-warp races replaced by a synchronous sweep.  See
 * pmmc_heatmap_sweep in include/pmmc_hip.h for the exact definition. */
typedef struct {
  double T0, dt, improvement, small_energy;
  float tjs, tjc, step, step_decay, sphere_radius;
  int32_t inner;           /* proposals per proposer per iteration (512) */
  int32_t proposers;       /* proposers per bead (64 * replicas)         */
  int32_t fails_threshold;
  int32_t max_iters;
  int32_t use2d;
} orc_sweep_schedule;

/* tie_rel > 0 counts near ties: decisions that an FP32 evaluation of the
 * scores with relative error tie_rel could take the other way (a proposal's
 * downhill or Metropolis test, the argmin over proposers), and FP64 ties of
 * the candidate-vs-snapshot test and of the stop test.  st->near_ties counts
 * iterations with a near tie, st->first_near_tie is the first one (0-based,
 * -1 if none); tie_beads [n] (nullable) flags the beads with a near tie in
 * that iteration (every bead for a stop-test tie).  e_trace (nullable,
 * max_iters + 1 entries): the full score before the first iteration and
 * after each one.  Returns the number of iterations. */
int orc_heatmap_sweep(const orc_heatmap *P, const orc_sweep_schedule *s, uint64_t seed,
                      uint32_t segment, float *pos, orc_stats *st, double tie_rel,
                      uint8_t *tie_beads, double *e_trace);
/* index of the lowest of n scores, ties to the lowest index */
int orc_sweep_argmin(const double *scores, int n);

/* ================= arc marking ================= */
/* ParallelMarkArcs.cu:26-64 / InteractionArcs.hpp:226-235: last anchor j with
 * length>1 containing the endpoint, else -1. */
void orc_mark_arcs(const int32_t *anc_s, const int32_t *anc_e, int32_t A,
                   const int32_t *arc_s, const int32_t *arc_e, int32_t R,
                   int32_t *out_s, int32_t *out_e);

/* ---- stress refinement (stress.c): LooperSolver::stressRefineAnchors'
 * descent, LooperSolver.hpp:5762-5910, float arithmetic as the reference ---- */
typedef struct {
  int iterations; /* position updates made */
  int stop;       /* 0 max_iter, 1 converged, 2 step below 1e-6, 3 nothing to do */
  float lr;
  double total, backbone, pairs, angular; /* last evaluated positions */
} orc_stress_stats;

void orc_stress_gradient(int n, const float *x, const float *bb, int np, const int *pi,
                         const int *pj, const float *pt, float ks, float kq, float ka, float *G,
                         double *e_back, double *e_hic, double *e_ang);
void orc_stress_refine(int n, float *x, const float *bb, int np, const int *pi, const int *pj,
                       const float *pt, float ks, float kq, float ka, int max_iter, float lr,
                       orc_stress_stats *st);
double orc_stress_energy64(int n, const float *x, const float *bb, int np, const int *pi,
                           const int *pj, const float *pt, double ks, double kq, double ka);
uint64_t orc_acosf_sum(uint32_t first, uint32_t count, int threads);

#ifdef __cplusplus
}
#endif
#endif
