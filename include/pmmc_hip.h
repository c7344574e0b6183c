/*
 * pmmc_hip.h -- C ABI of the MI355X-native pMMC annealing hot path.
 *
 * Pure C (stdint.h only): host C++ that includes this header compiles with
 * g++ and never sees a HIP type.  All pointers are caller-owned HOST memory;
 * device memory is owned by the opaque pmmc_ctx (grow-only arenas).
 *
 * Drop-in boundary (reference entry points this ABI replaces):
 *   pmmc_heatmap_anneal  <- float LooperSolver::ParallelMonteCarloHeatmap(float)
 *                           LooperSolver.hpp:469, ParallelMonteCarloHeatmap.cu:365
 *   pmmc_heatmap_sweep   <- MonteCarloHeatmapKernel's algorithm,
 *                           ParallelMonteCarloHeatmap.cu:217-363 (race-free form)
 *   pmmc_arcs_anneal     <- float LooperSolver::parallelMonteCarloArcs(float)
 *                           LooperSolver.hpp:682, ParallelMonteCarloArcs.cu:680
 *   pmmc_smooth_anneal   <- float LooperSolver::ParallelMonteCarloSmooth(float, ull)
 *                           LooperSolver.hpp:482, ParallelMonteCarloSmooth.cu:357
 *   pmmc_mark_arcs       <- kernel<<<256,1024>>> of InteractionArcs::parallelMarkArcs
 *                           ParallelMarkArcs.cu:36-64, :69-148
 *   pmmc_group_arcs      <- host grouping of InteractionArcs::parallelMarkArcs
 *                           ParallelMarkArcs.cu:150-243
 *   pmmc_set_gpu_seed    <- void setGpuSeed(unsigned) platform.h:106,
 *                           ParallelMonteCarloHeatmap.cu:37
 *   pmmc_singleton_heatmap <- LooperSolver::createSingletonHeatmap
 *                           LooperSolver.hpp:2517-2749 (records parsed)
 *   pmmc_ctx_create/destroy/reserve <- GpuResourceCache::init/cleanup and its
 *                           grow-only buffers, LooperSolver.hpp:54-106,
 *                           ParallelMonteCarloArcs.cu:84-213
 * Test hook (no reference counterpart, needed to check per-move ΔE on injected
 * proposals):  pmmc_delta_energy.
 * Batch / serve forms (no reference counterpart; the batched and served
 * annealing forms):  pmmc_anneal_batch, pmmc_server_*.
 * Diagnostics (no reference counterpart): pmmc_probe_move_pattern.
 *
 * Errors: every entry point returns pmmc_status; nothing calls exit().  The
 * message of the last failure is available from pmmc_last_error(ctx).
 */
#ifndef PMMC_HIP_H
#define PMMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PMMC_OK = 0,
  PMMC_ERR_INVALID = 1,   /* bad argument, size or layout                   */
  PMMC_ERR_HIP = 2,       /* a HIP runtime call failed                      */
  PMMC_ERR_NOMEM = 3,     /* device allocation failed / budget exceeded     */
  PMMC_ERR_NODEVICE = 4,  /* no usable gfx950 device                        */
  PMMC_ERR_BUSY = 5       /* server has no free slot / memory for the job   */
} pmmc_status;

typedef enum { PMMC_LEVEL_HEATMAP = 1, PMMC_LEVEL_ARCS = 2, PMMC_LEVEL_SMOOTH = 3 } pmmc_level;

/* Stop rules of the milestone test (see pmmc_schedule). */
typedef enum {
  PMMC_STOP_FIRST_FAIL = 0,   /* CPU annealers: stop at the first failing milestone
                                 (LooperSolver.hpp:1543-1548, :4775-4779, :5027-5031) */
  PMMC_STOP_MISSES_RESET = 1, /* GPU arcs: misses reset on improvement
                                 (ParallelMonteCarloArcs.cu:636-655)                 */
  PMMC_STOP_MISSES_ACCUM = 2  /* GPU heatmap / smooth: misses never reset
                                 (ParallelMonteCarloHeatmap.cu:343-352,
                                  ParallelMonteCarloSmooth.cu:320-335)               */
} pmmc_stop_rule;

typedef enum {
  PMMC_STOP_RUNNING = 0, PMMC_STOP_RULE = 1, PMMC_STOP_SMALL = 2, PMMC_STOP_CAP = 3,
  PMMC_STOP_GUARD = 4, PMMC_STOP_NOFREE = 5, PMMC_STOP_NONFINITE = 6
} pmmc_stop_reason;

/* Schedule of one Metropolis chain.  One chain performs blocks of `block`
 * moves.  Move m (0-based, counted) of chain c:
 *   bead p  = free[ pick(B, n_free) ], B = Philox(key, {m>>2, c>>pick_shift, segment,
 *             level<<28|1})[m&3]  (pick_shift = pmmc_pick_shift(level, n): 6, the
 *             64 chains of one wavefront share bead picks, or 0 for heatmaps of
 *             33..64 beads, whose register-resident kernels draw them per
 *             chain; displacement and acceptance are always per chain)
 *   {dx,dy,dz,U} = Philox(key, {m, c, segment, level<<28|0}), d = (2*u01-1)*step
 *   ΔE = local(p, x_p + d) - local(p, x_p)       (FP32, level semantics)
 *   E' = E + ΔE                                   (FP64 running total)
 *   accept if E' <= E (E' < E when strict_accept), else if T > 0 (and E > 0 when
 *   positive_guard): U < tjs * exp(-tjc * (E'/E) / T)
 *   T *= dt after every move (or once per block when cool_per_block)
 * After each block: step *= step_decay; when (iters % milestone_period) < block
 * a milestone runs: E is re-summed over the current positions (FP32 terms,
 * FP64 sum) and the stop rule is applied with
 *   no_improvement = (E > improvement*ms && successes < min_successes)
 *                    || E < small_energy || (ratio_stop > 0 && ms > 1e-10 && E/ms > ratio_stop)
 * Chains stop when the rule fires, E is non-finite, or iters >= max_moves
 * (checked before each block, as the reference GPU kernels do).
 * Presets reproducing each reference annealer: pmmc_schedule_preset().      */
typedef struct {
  double T0;              /* maxTemp*                                           */
  double dt;              /* dtTemp*                                            */
  double improvement;     /* MCstopConditionImprovement*                        */
  double small_energy;    /* 1e-6 (CPU heatmap/smooth) / 1e-5 (arcs) / 1e-4 (GPU heatmap) */
  double ratio_stop;      /* 0.9999 for arcs, 0 = off                           */
  double guard_energy;    /* skip annealing when E0 < guard or NaN (heatmap 1e-6) */
  float tjs;              /* tempJumpScale*                                     */
  float tjc;              /* tempJumpCoef*                                      */
  float step;             /* displacement half width, already scaled            */
  float step_decay;       /* per block (0.95 for the GPU heatmap schedule)      */
  float sphere_radius;    /* Settings::sphere_radius, 0 = off                   */
  int32_t block;          /* moves per block                                    */
  int32_t milestone_period;
  int32_t min_successes;  /* MCstopConditionMinSuccesses*                       */
  int32_t fails_threshold;/* [cuda] milestone_fails                             */
  int32_t stop_rule;      /* pmmc_stop_rule                                     */
  int32_t strict_accept;
  int32_t positive_guard;
  int32_t cool_per_block;
  int32_t use2d;
  int32_t resum;          /* K > 0: replace E by its FP64 re-sum at every K-th
                             milestone (before the stop test); 0: never         */
  int64_t max_moves;      /* hard cap per chain                                 */
} pmmc_schedule;

/* pick_shift of the bead-pick stream above for a problem of `level` with n
 * beads: 0 (picks per chain) for heatmaps of 33..64 beads, 6 otherwise. */
int32_t pmmc_pick_shift(int32_t level, int32_t n);

/* Which reference annealer a preset reproduces. */
typedef enum { PMMC_PRESET_CPU = 0, PMMC_PRESET_GPU = 1 } pmmc_preset;

/* Reference settings consumed by the presets ([cuda] and [simulation_*] INI
 * sections, Settings.hpp:649-654, :865-941, defaults :350-515). */
typedef struct {
  double max_temp, dt_temp, improvement;
  float tjs, tjc;
  int32_t min_successes, steps;   /* stop_condition_successes / stop_condition_steps */
  int32_t milestone_fails;        /* [cuda] milestone_fails                            */
  float sphere_radius;
  int32_t use2d;
} pmmc_level_settings;

/* Fill *out with the schedule that reproduces the reference annealer of
 * `level` under `preset` (CPU: LooperSolver::MonteCarlo*; GPU: the .cu
 * kernels), for a caller step size `step_size` (the value the reference level
 * driver passes).  Returns PMMC_ERR_INVALID for an unknown level/preset. */
pmmc_status pmmc_schedule_preset(int level, int preset, const pmmc_level_settings *s,
                                 float step_size, pmmc_schedule *out);

/* ---------------- problems (host memory, caller owned) ---------------- */

/* Heatmap level: LooperSolver state read by ParallelMonteCarloHeatmap. */
typedef struct {
  int32_t n;                 /* active region size == heatmap size          */
  const float *pos;          /* [n*3] initial xyz                           */
  const uint8_t *fixed;      /* [n] or NULL                                  */
  const float *dist;         /* [n*n] heatmap_dist row-major                */
  int32_t diagonal;          /* heatmap_dist.diagonal_size                  */
  const int32_t *boundaries; /* heatmap_chromosome_boundaries or NULL       */
  int32_t n_boundaries;
  float skip_below;          /* 1e-6 (CPU, LooperSolver.hpp:3604); 1e-3 (GPU, .cu:176) */
} pmmc_heatmap_problem;

/* Arc level: state read by parallelMonteCarloArcs. */
typedef struct {
  int32_t n;
  const float *pos;          /* [n*3]                                       */
  const uint8_t *fixed;      /* [n] or NULL                                  */
  const float *exp_dist;     /* [n*n] heatmap_exp_dist_anchor, -1 = no arc  */
  int32_t n_loops;
  const int32_t *loop_pairs; /* [2*n_loops] active-region indices            */
  const float *loop_params;  /* [2*n_loops] {stiffness, eq_distance}         */
  float k_stretch, k_squeeze;/* springConstantStretchArcs / SqueezeArcs     */
} pmmc_arcs_problem;

/* Smooth level: state read by ParallelMonteCarloSmooth. */
typedef struct {
  int32_t n;
  const float *pos;          /* [n*3]                                       */
  const uint8_t *fixed;      /* [n] or NULL                                  */
  const float *dist_to_next; /* [n]                                          */
  float k_stretch, k_squeeze, k_angular; /* [springs]                        */
  float w_dist, w_angle;     /* weightDistSmooth, weightAngleSmooth          */
} pmmc_smooth_problem;

/* How many chains and which random stream. */
typedef struct {
  uint64_t seed;             /* Philox key                                   */
  uint32_t segment;          /* Philox counter word 2 (problem / segment id) */
  uint32_t chain0;           /* first chain id                               */
  int32_t n_chains;          /* 0 = library default for the level            */
} pmmc_run;

typedef struct {
  double energy;             /* best chain: final running total (FP64)        */
  double energy_resummed;    /* best chain: level return value re-summed in FP64:
                                heatmap calcScoreHeatmapActiveRegion(),
                                arcs running score, smooth calcScoreStructureSmooth(true,true) */
  double best_energy;        /* lowest running total seen by the best chain    */
  int64_t moves;             /* counted moves, summed over all chains          */
  int64_t accepted;          /* accepted moves, summed over all chains         */
  int64_t best_chain_moves;  /* moves of the returned chain                    */
  int64_t best_chain_accepted;
  int32_t best_chain;        /* chain id of the returned chain                 */
  int32_t n_chains;
  int32_t milestones;        /* of the returned chain                          */
  int32_t stop_reason;       /* of the returned chain (pmmc_stop_reason)      */
  double kernel_ms;          /* device time of the annealing launches          */
} pmmc_anneal_stats;

/* ---------------- context ---------------- */
typedef struct pmmc_ctx pmmc_ctx;

pmmc_status pmmc_ctx_create(int device, pmmc_ctx **out);
void pmmc_ctx_destroy(pmmc_ctx *ctx);
/* Grow the device arena to at least `bytes` (HBM-budget admission). */
pmmc_status pmmc_ctx_reserve(pmmc_ctx *ctx, size_t bytes);
const char *pmmc_last_error(const pmmc_ctx *ctx);
/* Number of compute units of the context's device. */
int pmmc_ctx_cu_count(const pmmc_ctx *ctx);

/* setGpuSeed: default seed used when a run passes seed == 0. */
void pmmc_set_gpu_seed(pmmc_ctx *ctx, uint64_t seed);

/* Chains an annealing call runs for a problem of `level` with n beads when
 * run.n_chains = 0 (no reference counterpart: the reference sizes its launch
 * from the SM count, ParallelMonteCarloArcs.cu:700-751): whole rounds of the
 * level kernel's resident wavefronts (64 chains each) per compute unit, at
 * least 16 wavefronts per CU, capped so that the chains' positions take at
 * most a third of the device's HBM.  symmetric (heatmap level): whether the
 * heatmap's weight tables are symmetric, which selects the kernel for 33..64
 * beads.  Returns 0 for a null context, an unknown level or n < 1. */
int32_t pmmc_default_chains(const pmmc_ctx *ctx, int32_t level, int32_t n, int32_t symmetric);

/* ---------------- plain (one problem per call) ---------------- */
/* Anneal n_chains chains from problem->pos, write the positions of the chain
 * with the lowest final running energy (ties: lowest chain id) to pos_out
 * [n*3].  For the smooth level only non-fixed beads are written back, as
 * ParallelMonteCarloSmooth.cu:485-491 does; pos_out must hold the input
 * positions for the fixed beads. */
pmmc_status pmmc_heatmap_anneal(pmmc_ctx *ctx, const pmmc_heatmap_problem *p,
                                const pmmc_schedule *s, const pmmc_run *run,
                                float *pos_out, pmmc_anneal_stats *stats);
pmmc_status pmmc_arcs_anneal(pmmc_ctx *ctx, const pmmc_arcs_problem *p,
                             const pmmc_schedule *s, const pmmc_run *run,
                             float *pos_out, pmmc_anneal_stats *stats);
pmmc_status pmmc_smooth_anneal(pmmc_ctx *ctx, const pmmc_smooth_problem *p,
                               const pmmc_schedule *s, const pmmc_run *run,
                               float *pos_out, pmmc_anneal_stats *stats);

/* ---------------- plain, allocation-free ---------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): the plain annealing call of `level` (same problem
 * structs, kernels, results and write-back as pmmc_heatmap_anneal /
 * pmmc_arcs_anneal / pmmc_smooth_anneal, bit for bit) with every device
 * buffer carved from the library's static per-device scratch array of
 * 64 MiB; it calls no hipMalloc, hipFree or other allocating routine.
 * run->n_chains = 0 takes the library default reduced to the most whole
 * wavefronts (64 chains) that fit; a problem or an explicit chain count that
 * needs more than 64 MiB is rejected with PMMC_ERR_NOMEM (the message gives
 * the bytes needed).  Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (other threads or contexts) race. */
pmmc_status pmmc_anneal_noalloc(pmmc_ctx *ctx, int32_t level, const void *problem,
                                const pmmc_schedule *s, const pmmc_run *run, float *pos_out,
                                pmmc_anneal_stats *stats);

/* ---------------- heatmap sweep (reference GPU heatmap kernel) ---------------- */
/* Each iteration: every non-fixed bead b gets `proposers` independent
 * proposers; proposer r runs `inner` Metropolis proposals on bead b against a
 * snapshot of all other beads (score = local(b) + sphere, GPU skip threshold),
 * starting from the snapshot position and from the previous iteration's best
 * score (the initial full score in the first iteration).  The proposer with the
 * lowest final score (ties: lowest r) replaces bead b when its local score is
 * below the bead's current one.  All beads update together.  Then T *= dt,
 * step *= step_decay and the full score is re-summed in FP64; misses counts
 * iterations with E > improvement * previous (never reset); stop when
 * misses >= fails, E < small_energy, E is non-finite, or max_iters is reached.
 * Proposer (b, r) uses the per-chain stream with chain = b*proposers + r and
 * move = iteration*inner + k at level 4. */
typedef struct {
  double T0, dt, improvement, small_energy;
  float tjs, tjc, step, step_decay, sphere_radius;
  int32_t inner;             /* 512 in the reference                           */
  int32_t proposers;         /* multiple of 64                                 */
  int32_t fails_threshold;
  int32_t max_iters;
  int32_t use2d;
} pmmc_sweep_schedule;

pmmc_status pmmc_heatmap_sweep(pmmc_ctx *ctx, const pmmc_heatmap_problem *p,
                               const pmmc_sweep_schedule *s, uint64_t seed, uint32_t segment,
                               float *pos_inout, pmmc_anneal_stats *stats);

/* ---------------- singleton heatmap (heatmap input construction) ---------------- */
/* <- LooperSolver::createSingletonHeatmap (LooperSolver.hpp:2517-2749): the
 * contact heatmap between the clusters of the current level, from singleton
 * PET records.  The reference reads the records from text files
 * (:2590-2626); here they come parsed, each chromosome given as an index into
 * the level's chromosome list (< 0 or >= n_chr: a chromosome that is not
 * processed, counted in bad_chr).  Bins per chromosome: 0, the midpoints of
 * consecutive clusters ((end_i + start_i+1) / 2, integer), 1e9; a position
 * falls in the first bin whose closed range holds it (:2652-2663); records
 * whose start1 or end2 lies outside [region_start, region_end] are skipped
 * when region_end > 0 (:2633-2637); negative start1 / start2 count as 0;
 * records with both ends in one bin are counted in `diagonal` and skipped.
 * Scores are summed as 64-bit integers on the device (the reference sums them
 * in float: the same value while a cell stays below 2^24).  Then, as the
 * reference: each long-range arc adds (float)(long_scale * score^long_power)
 * twice to cell [bin of start][bin of end] (:2701-2737, as written); the
 * first and last break of each chromosome become its lowest start and
 * highest end among the added records (:2740-2746); bin lengths exclude the
 * chromosome's centromere (getGenomicLengthExcludingCentromere, :5916-5935)
 * and are in Mb; the upper triangle divided by len_i * len_j is mirrored
 * (:2773-2779); the band |i - j| < diag is cleared (Heatmap::clearDiagonal).
 * A record or long arc outside every bin is an error here (the reference
 * exits, :2665-2668, or indexes bin -1). */
typedef struct {
  int32_t n_chr;
  const int32_t *cluster_offsets;  /* [n_chr + 1]: clusters of chromosome c   */
  const int32_t *cluster_start;    /* genomic start / end of the level's       */
  const int32_t *cluster_end;      /* clusters, chromosome by chromosome       */
  int64_t n_records;
  const int32_t *chr1, *start1, *end1;  /* first end of each record           */
  const int32_t *chr2, *start2, *end2;  /* second end                          */
  const int32_t *score;
  int32_t region_start, region_end;
  const int32_t *centromere;   /* [3 * n_chr]: has, start, end per chromosome
                                  (the first region of the chromosome), or
                                  NULL: no centromere data                   */
  int32_t n_long;              /* long-range arcs, each chromosome's in the  */
  const int32_t *long_chr;     /* reference's order                          */
  const int32_t *long_start, *long_end, *long_score;
  float long_scale, long_power;  /* Settings::longPETClustersEffectScale / Power */
  int32_t diag;
} pmmc_singleton_input;

typedef struct {
  int64_t added, diagonal, bad_chr;  /* records added, on one bin, on other chromosomes */
  double kernel_ms;                  /* device time of the binning launches          */
} pmmc_singleton_counts;

/* heat: N x N row-major, N = cluster_offsets[n_chr]; counts nullable. */
pmmc_status pmmc_singleton_heatmap(pmmc_ctx *ctx, const pmmc_singleton_input *in, float *heat,
                                   pmmc_singleton_counts *counts);

/* ---------------- bandwidth probe (diagnostics) ---------------- */
/* No reference counterpart: measures what the GPU sustains for the move
 * kernels' own access patterns, the ceiling their throughput is read against
 * besides the nominal HBM peak.  Per move of a 64-lane wavefront: `reads`
 * random windows of `chunk_beads` consecutive beads (768 B per bead, laid
 * out as the level's kernel lays it: chunk_beads 1 with reads <= 8 -- the arc
 * move, three 256-byte SoA rows per bead -- or 5 with reads 1 -- the smooth
 * move, each lane's 12 bytes together), then the first window's first bead is
 * written back by no lane (write_mode 0), about half of the lanes (1), every
 * lane (2) or one lane in every other move (3); two moves in flight per
 * wavefront; `waves` wavefronts of
 * `moves_per_wave` (even) moves over a buffer of `bytes` of context-owned
 * device memory.  Also a streaming read of that buffer and a conflict-free
 * ds_read_b128 loop over 64 KB of LDS per workgroup. */
typedef struct {
  double moves_per_s;       /* lane-moves per second of the pattern        */
  double read_gbps;         /* bytes of the pattern read per second / 1e9  */
  double write_gbps;        /* bytes of the pattern written / s / 1e9      */
  double stream_read_gbps;  /* streaming read of the buffer                */
  double lds_gbps;          /* aggregate LDS read bandwidth                */
  double ms;                /* duration of the pattern launch              */
} pmmc_probe_result;

pmmc_status pmmc_probe_move_pattern(pmmc_ctx *ctx, size_t bytes, int32_t waves,
                                    int32_t moves_per_wave, int32_t reads, int32_t chunk_beads,
                                    int32_t write_mode, pmmc_probe_result *out);

/* ---------------- per-move ΔE test hook ---------------- */
/* Evaluate ΔE (FP32, the kernels' arithmetic) for n_prop injected proposals
 * (beads[i], disp[3*i..3*i+2]) against the fixed state `pos` of the problem.
 * problem points to the level's problem struct; sphere_radius as in the
 * schedule. */
pmmc_status pmmc_delta_energy(pmmc_ctx *ctx, int level, const void *problem,
                              float sphere_radius, int32_t n_prop, const int32_t *beads,
                              const float *disp, float *delta_out);

/* ---------------- arc marking ---------------- */
/* For each raw arc i: out_start[i] = last anchor j with length_j > 1 that
 * contains arc_start[i] (-1 if none), likewise out_end.  Anchor length is 0
 * when start == end == 0, else end - start + 1 (Anchor.hpp:81-87). */
pmmc_status pmmc_mark_arcs(pmmc_ctx *ctx, const int32_t *anchor_start,
                           const int32_t *anchor_end, int32_t n_anchors,
                           const int32_t *arc_start, const int32_t *arc_end, int32_t n_arcs,
                           int32_t *out_start, int32_t *out_end);

/* One emitted arc (InteractionArc fields, InteractionArc.hpp:86-92). */
typedef struct {
  int32_t start, end, genomic_start, genomic_end, score, eff_score, factor;
} pmmc_arc;

/* Whole InteractionArcs::parallelMarkArcs for n_chr chromosomes in order:
 * sort raw arcs (start, end, factor), mark on the GPU, group by start anchor
 * and merge per factor exactly as the reference host loop does, including the
 * std::unordered_map emission order and the cross-chromosome last_start.
 * raw: [4 * total] (genomic start, genomic end, score, factor) per chromosome
 * slice raw_off[c]..raw_off[c+1]; anchors likewise via anc_off.  out must
 * hold 2 * total arcs; out_off[n_chr+1] receives the per-chromosome slices;
 * nonmatching[n_chr] (nullable) the skipped-arc counts. */
pmmc_status pmmc_group_arcs(pmmc_ctx *ctx, int32_t n_chr, const int32_t *anc_off,
                            const int32_t *anc_start, const int32_t *anc_end,
                            const int32_t *raw_off, const int32_t *raw, pmmc_arc *out,
                            int32_t *out_off, int32_t *nonmatching);

/* ---------------- batched form ---------------- */
typedef struct {
  int32_t level;             /* pmmc_level                                     */
  const void *problem;       /* pmmc_{heatmap,arcs,smooth}_problem*            */
  pmmc_schedule schedule;
  pmmc_run run;
  float *pos_out;            /* [n*3] result (smooth: non-fixed beads only)     */
  pmmc_anneal_stats *stats;  /* nullable                                        */
} pmmc_job;

typedef struct {
  int32_t blocks_per_step;   /* blocks of moves per step launch (0 = autotune);
                                -1: one launch per level group that runs every
                                chain to its end, each workgroup keeping its
                                ensemble until its last chain stops (the
                                workgroup dispatcher starts the ensembles
                                beyond the resident ones as slots free up: a
                                persistent pool with the dispatcher as its
                                queue); no steps, graph or tuning           */
  int32_t steps_per_graph;   /* step launches captured per hipGraph (0 = auto)  */
  int32_t use_graph;         /* 1: replay captured step loops                  */
  int32_t chains_per_wave;   /* arc level: 64 (a lane per chain) or 32 (a lane
                                pair per chain: twice the wavefronts for the
                                same chains); 0 = autotune.  Results do not
                                depend on it.                                  */
  int32_t tuned;             /* out: 1 when a field above was chosen by timing  */
} pmmc_batch_opts;

/* Run all jobs concurrently (one device-resident chain pool, step launches
 * replayed from a captured hipGraph until every chain has stopped).  Fields
 * of *opts left 0 are chosen by an autotuner that times the first steps of
 * the run itself (each candidate launch shape runs one short step of about
 * 1,024 moves per chain; the chains advance meanwhile, so no work is
 * repeated): chains_per_wave by the higher moves/s, blocks_per_step for
 * about 20 ms of device time per step, steps_per_graph for about 100 ms per
 * graph replay; the values used are written back to *opts.  Jobs that name
 * the same level and problem struct as an earlier job share its host
 * preparation (the tables are built once and copied). */
pmmc_status pmmc_anneal_batch(pmmc_ctx *ctx, const pmmc_job *jobs, int32_t n_jobs,
                              pmmc_batch_opts *opts);

/* ---------------- served form (continuous batching) ---------------- */
typedef struct pmmc_server pmmc_server;

typedef struct {
  size_t hbm_budget;         /* bytes of chain state + problem data admitted;
                                allocated on the device once, at creation      */
  int32_t max_jobs;          /* resident job slots                              */
  int32_t blocks_per_step;   /* blocks of moves per chain per step launch; 0 =
                                paced per launch group (kernel and size class)
                                to about launch_ms of device time per launch   */
  float launch_ms;           /* pacing target per group launch; 0 = 12 ms      */
  int32_t streams;           /* streams a step's group launches are spread
                                over (1-16, all at the default priority);
                                0 = 4                                          */
} pmmc_server_opts;

pmmc_status pmmc_server_create(pmmc_ctx *ctx, const pmmc_server_opts *opts, pmmc_server **out);
void pmmc_server_destroy(pmmc_server *srv);
/* Queue a job (inputs are copied).  *job_id identifies it in poll. */
pmmc_status pmmc_server_submit(pmmc_server *srv, const pmmc_job *job, uint64_t *job_id);
/* Queue n_jobs jobs (no reference counterpart: for requests of several
 * solves, e.g. an ensemble of seeds on one problem).  Jobs that name the same
 * level and problem struct as an earlier job of the list share its host
 * preparation (the tables are built once and copied); each job's results are
 * those pmmc_server_submit of it alone gives.  job_ids[k] identifies job k
 * (nullable).  On error nothing is queued. */
pmmc_status pmmc_server_submit_many(pmmc_server *srv, const pmmc_job *jobs, int32_t n_jobs,
                                    uint64_t *job_ids);
/* Admit queued jobs that fit the budget, run one launch per launch group over
 * every resident chain, retire finished jobs (results written to the job's
 * pos_out / stats).  *n_done = jobs finished in this step; *n_resident and
 * *n_queued describe the pool afterwards (nullable).  A job whose upload
 * fails at admission is dropped (it never reports done) and the step returns
 * that error; the server stays usable. */
pmmc_status pmmc_server_step(pmmc_server *srv, int32_t *n_done, int32_t *n_resident,
                             int32_t *n_queued);
/* 1 when job_id has finished (its outputs are written). */
int pmmc_server_done(const pmmc_server *srv, uint64_t job_id);

/* ---------------- equidistant model (spline) ---------------- */
/* Number of outputs the reference allocates for [pos_start, pos_stop] at
 * resolution_bp: (pos_stop - pos_start + 1) / resolution_bp + 1
 * (HierarchicalChromosome.cu:1268). */
int64_t pmmc_spline_count(int64_t pos_start, int64_t pos_stop, int32_t resolution_bp);

/* Catmull-Rom positions of one chromosome's bead chain at genomic positions
 * pos_start + k*resolution_bp, k < n_out <- HierarchicalChromosome::gpuHelper
 * (HierarchicalChromosome.cu:1252-1449) and its kernel
 * parallelFind3DSmoothPosition (:420-464).  genomic_pos [n_beads] sorted
 * (the level order after HierarchicalChromosome::levelDown, :232-257), xyz
 * [3*n_beads]; first_end / last_start = end of the first bead, start of the
 * last.  Every output is written (the reference kernel leaves outputs beyond
 * one grid unwritten); arithmetic is the reference's CPU path's, bit for bit. */
pmmc_status pmmc_spline_positions(pmmc_ctx *ctx, int32_t n_beads, const int32_t *genomic_pos,
                                  const float *xyz, int32_t first_end, int32_t last_start,
                                  int64_t pos_start, int32_t resolution_bp, int64_t n_out,
                                  float *out_xyz);

/* ---------------- stress refinement of anchor positions ---------------- */
/* One problem of LooperSolver::stressRefineAnchors (LooperSolver.hpp:5638-5915)
 * after its pair list is built: gradient descent on
 *   E = sum_i k ((|x_i - x_i+1| - b_i) / b_i)^2      springs with b_i > 1e-6,
 *                                                     k = k_stretch when
 *                                                     stretched, else k_squeeze
 *     + sum_p ((|x_i - x_j| - t_p) / t_p)^2          pairs (i, j, t_p)
 *     + sum_{i=1..n-2} k_angular theta_i^3            bending angle at bead i
 * with the bending gradient by forward differences (step 1e-3).  Each
 * iteration halves the step if the energy rose (stopping, before moving, once
 * it falls below 1e-6), moves x -= step * gradient, and stops after iteration
 * 21 once the energy fell by less than 1e-7 max(1, E).  The reference builds
 * the pairs from the chromosome's arcs and long arcs (pmmc_tree_stress_refine
 * does that from a tree); here they are given. */
typedef struct {
  int32_t n;                   /* beads, in chain (genomic) order                */
  const float *pos;            /* [3n] start positions                           */
  const float *backbone;       /* [n-1] spring target to the next bead; <= 1e-6:
                                  no spring                                      */
  int32_t n_pairs;
  const int32_t *pair_i;       /* [n_pairs] bead indices, i != j                 */
  const int32_t *pair_j;
  const float *pair_target;    /* [n_pairs] target distance                      */
  float k_stretch, k_squeeze, k_angular;
} pmmc_stress_problem;

typedef struct {
  int32_t max_iter;       /* stressRefineIterations (default 0: nothing to do) */
  float lr;               /* stressRefineLR (default 0.05)                      */
  int32_t in_order_sums;  /* 1: sum every iteration's energy terms one at a time,
                             in the reference's order (slower).  0: a tree sum,
                             falling back to the in-order sums only when a
                             decision lies within their rounding-error bound;
                             the results are the same either way.             */
} pmmc_stress_opts;

void pmmc_stress_opts_default(pmmc_stress_opts *o);

typedef enum {
  PMMC_STRESS_MAX_ITER = 0,   /* ran max_iter iterations                        */
  PMMC_STRESS_CONVERGED = 1,  /* energy change below 1e-7 max(1, E)             */
  PMMC_STRESS_LR_FLOOR = 2,   /* step halved below 1e-6                         */
  PMMC_STRESS_NOTHING = 3     /* max_iter <= 0 or fewer than 2 beads            */
} pmmc_stress_stop;

typedef struct {
  int32_t iterations;     /* position updates made                             */
  int32_t stop;           /* pmmc_stress_stop                                  */
  float lr;               /* step at the end                                   */
  int32_t exact_sums;     /* iterations whose decision used the in-order sums  */
  double total;           /* energy of the last evaluated positions (in-order  */
  double backbone;        /*   sums, as the reference prints them): total =    */
  double pairs;           /*   (backbone + pairs) + angular                    */
  double angular;
} pmmc_stress_stats;

/* Refines `count` problems in one launch (one workgroup each); pos_out[c]
 * [3n] receives problem c's positions (it may alias probs[c].pos); stats
 * nullable.  Results are the reference's bit for bit. */
pmmc_status pmmc_stress_refine(pmmc_ctx *ctx, int32_t count, const pmmc_stress_problem *probs,
                               const pmmc_stress_opts *opts, float *const *pos_out,
                               pmmc_stress_stats *stats);

/* Diagnostics (no reference counterpart): the device's arc cosine (the
 * kernels' replacement for the reference's libm acosf) over the floats whose
 * bit patterns are first_bits .. first_bits + count - 1, folded into
 * *sum = sum over inputs of splitmix64((bits << 32) | result bits) mod 2^64,
 * so a host computing the same sum with libm's acosf can check every input. */
pmmc_status pmmc_probe_acosf(pmmc_ctx *ctx, uint32_t first_bits, uint32_t count, uint64_t *sum);

#ifdef __cplusplus
}
#endif
#endif
