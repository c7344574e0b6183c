/*
 * pmmc_levels.h -- C ABI of the host-side rows around the annealing kernels:
 * the cluster tree and bead indexing, sub-anchor densification, the arc and
 * heatmap expected-distance inputs, and the level drivers that run the
 * annealers over a tree.
 *
 * Pure C (stdint.h only), caller-owned host memory, pmmc_status returns as in
 * pmmc_hip.h.  Host code only: everything here runs on the CPU except the
 * level drivers, which launch the annealing kernels through pmmc_hip.h.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference's src/ directory):
 *   pmmc_tree_create        LooperSolver::createTreeGenome / createTreeChromosome
 *                           LooperSolver.hpp:2272-2438, findGaps :2119-2180,
 *                           findSplit :2185-2269 (arcs shifted to global
 *                           cluster indices :2319-2329)
 *   pmmc_tree_level         LooperSolver::setLevel / levelDown :2441-2452,
 *                           :3087-3104 (children in stored order, no sort)
 *   pmmc_tree_densify       LooperSolver::densifyActiveRegion :3836-3898
 *   pmmc_tree_set_dist_to_next  dist_to_next / noise_size of
 *                           reconstructClusterArcsDistances :4339-4354 with
 *                           genomicLengthToDistance :3900-3904
 *   pmmc_tree_expected_distances
 *                           LooperSolver::calcAnchorExpectedDistancesHeatmap
 *                           :5494-5627
 *   pmmc_freq_to_distance   LooperSolver::freqToDistance :3949-3967
 *   pmmc_distance_heatmap   LooperSolver::createDistanceHeatmap :3042-3085,
 *                           freqToDistanceHeatmap(Inter) :3934-3942
 *   pmmc_normalize_heatmap  LooperSolver::normalizeHeatmap :2996-3040
 *   pmmc_normalize_heatmap_diagonal_total
 *                           LooperSolver::normalizeHeatmapDiagonalTotal
 *                           :3146-3172
 *   pmmc_normalize_heatmap_inter
 *                           LooperSolver::normalizeHeatmapInter :3106-3144
 *   pmmc_tree_stress_refine LooperSolver::stressRefineAnchors :5638-5915
 *   pmmc_arcs_level (global_refine)
 *                           the global anchor refinement at the end of
 *                           reconstructClustersArcsDistances' chromosome
 *                           loop :4141-4243
 *
 * Errors: the message of the last failure on the calling thread is returned
 * by pmmc_host_last_error() (the reference prints and exits instead).
 */
#ifndef PMMC_LEVELS_H
#define PMMC_LEVELS_H

#include <stdint.h>

#include "pmmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Settings of the host-side rows; names follow Settings.hpp.  Defaults
 * (pmmc_level_params_default) are the reference's (Settings.hpp:395-461). */
typedef struct {
  float count_a;            /* countToDistA: < 0 selects the power-law branch  */
  float count_scale;        /* countToDistScale                                */
  float count_shift;        /* countToDistShift                                */
  float count_base;         /* countToDistBaseLevel                            */
  float heat_scale;         /* freqToDistHeatmapScale                          */
  float heat_power;         /* freqToDistHeatmapPower                          */
  float heat_scale_inter;   /* freqToDistHeatmapScaleInter                     */
  float heat_power_inter;   /* freqToDistHeatmapPowerInter                     */
  float heat_stretching;    /* heatmapDistanceHeatmapStretching                */
  float glen_base;          /* genomicLengthToDistBase                         */
  float glen_scale;         /* genomicLengthToDistScale                        */
  float glen_power;         /* genomicLengthToDistPower                        */
  int32_t loop_density;     /* loopDensity: beads inserted between anchors     */
  float noise_chr;          /* noiseCoefficientLevelChr                        */
  float noise_segment;      /* noiseCoefficientLevelSegment                    */
  float noise_anchor;       /* noiseCoefficientLevelAnchor                     */
  float noise_subanchor;    /* noiseCoefficientLevelSubanchor                  */
  float noise_anchor_small; /* noiseAnchorSmall                                */
  int32_t restarts_chr;     /* simulationStepsLevelChr                         */
  int32_t restarts_segment; /* simulationStepsLevelSegment                     */
  int32_t restarts_anchor;  /* simulationStepsLevelAnchor                      */
  int32_t restarts_subanchor; /* simulationStepsLevelSubanchor                 */
  int32_t use_anchor_heatmap;      /* useAnchorHeatmap                         */
  float anchor_heatmap_influence;  /* anchorHeatmapInfluence                   */
  int32_t use2d;            /* use2D                                           */
} pmmc_level_params;

void pmmc_level_params_default(pmmc_level_params *p);

/* Message of the last failed call of this header on the calling thread. */
const char *pmmc_host_last_error(void);

/* ---- expected-distance rules ---------------------------------------------- */

/* PET count -> expected distance (float, as the reference returns it).  The
 * reference's memo leaves count 1 uninitialised (LooperSolver.hpp:950-951);
 * here every count is computed. */
float pmmc_freq_to_distance(int32_t count, const pmmc_level_params *p);
/* genomic length (bp) -> distance, double as the reference computes it */
double pmmc_genomic_length_to_distance(int32_t length, const pmmc_level_params *p);

/* Distance heatmap of an n x n frequency heatmap: -1 inside the diagonal
 * band (|i-j| < diagonal_size), 0 where the frequency is below 1e-6, else
 * scale * f^power (the inter-chromosome rule when inter != 0, which the
 * reference uses for heatmap index 0); then values above
 * avg * heat_stretching are clamped.  Only the upper triangle of freq is
 * read (the output is symmetric).  out: n*n row-major.  *avg_out (nullable)
 * = the average the clamp used (Heatmap::getAvg, Heatmap.hpp:440-451). */
pmmc_status pmmc_distance_heatmap(const float *freq, int32_t n, int32_t diagonal_size,
                                  int32_t inter, const pmmc_level_params *p, float *out,
                                  float *avg_out);

/* Frequency-heatmap normalisations the level drivers apply between
 * createSingletonHeatmap (pmmc_singleton_heatmap, pmmc_hip.h) and
 * createDistanceHeatmap (LooperSolver.hpp:1095-1121, :1203-1215).  Float
 * arithmetic and summation order of the reference. */
/* Rows scaled to the average row sum (a row summing to <= 1e-12 becomes 0),
 * then each pair set to the mean of its two triangles.  in, out: n*n
 * row-major (may alias). */
pmmc_status pmmc_normalize_heatmap(const float *in, int32_t n, float *out);
/* Scaled in place so that the first non-zero diagonal (Heatmap::
 * getDiagonalSize) averages val (FP64 average); unchanged when that average
 * is not above 1e-12. */
pmmc_status pmmc_normalize_heatmap_diagonal_total(float *heat, int32_t n, float val);
/* Every cell times scale, then the intra-chromosome blocks (chromosome c has
 * sizes[c] consecutive bins) divided by it again; unchanged for one
 * chromosome. */
pmmc_status pmmc_normalize_heatmap_inter(float *heat, int32_t n, int32_t n_chr,
                                         const int32_t *sizes, float scale);

/* ---- cluster tree ---------------------------------------------------------- */

typedef struct pmmc_tree pmmc_tree;

/* One chromosome after arc marking (InteractionArcs::anchors / arcs /
 * long_arcs, InteractionArcs.hpp:151-170). */
typedef struct {
  const char *name;
  int32_t n_anchors;
  const int32_t *anchor_start;  /* [n_anchors] bp, sorted by start          */
  const int32_t *anchor_end;    /* [n_anchors] bp                           */
  const char *orientation;      /* [n_anchors] 'L' 'R' 'N', nullable        */
  int32_t n_arcs;
  const int32_t *arc_start;     /* [n_arcs] anchor index (chromosome-local) */
  const int32_t *arc_end;       /* [n_arcs]                                  */
  const int32_t *arc_score;     /* [n_arcs] PET count                        */
  int32_t n_long_arcs;          /* long-range arcs, genomic positions       */
  const int32_t *long_start;
  const int32_t *long_end;
  const int32_t *long_score;
} pmmc_chromosome;

/* One predefined segment (the reference's segment-split BED file, in file
 * order). */
typedef struct {
  const char *chr;
  int32_t start, end;
} pmmc_segment_region;

/* One cluster of the tree (Cluster.hpp:86-102).  Levels: 1 chromosome root,
 * 2 segment, 3 interaction block, 4 anchor, 5 sub-anchor bead. */
typedef struct {
  int32_t start, end, genomic_pos;
  int32_t base_start, base_end;
  int32_t level;
  int32_t parent;              /* -1 for roots and sub-anchor beads         */
  int32_t n_children;
  int32_t is_fixed;
  int32_t orientation;         /* 'L' 'R' 'N'                               */
  float pos[3];
  double dist_to_next;
} pmmc_cluster;

/* Builds the tree of every chromosome, in order: per chromosome its anchors,
 * then interleaved interaction blocks and segments, then its root.  With
 * n_regions > 0 the segment boundaries follow the predefined regions (the
 * path every pMMC run takes); with none, every interaction block is a
 * segment.  Fails where the reference calls error() or reads past a list. */
pmmc_status pmmc_tree_create(int32_t n_chr, const pmmc_chromosome *chrs, int32_t n_regions,
                             const pmmc_segment_region *regions, pmmc_tree **out);
void pmmc_tree_destroy(pmmc_tree *t);

int32_t pmmc_tree_size(const pmmc_tree *t);
int32_t pmmc_tree_n_chromosomes(const pmmc_tree *t);
pmmc_status pmmc_tree_get(const pmmc_tree *t, int32_t first, int32_t count, pmmc_cluster *out);
/* children of a cluster; *n = count even when it exceeds cap */
pmmc_status pmmc_tree_children(const pmmc_tree *t, int32_t cluster, int32_t *out, int32_t cap,
                               int32_t *n);
pmmc_status pmmc_tree_chromosome(const pmmc_tree *t, int32_t chr, int32_t *first_cluster,
                                 int32_t *root);
/* arcs of a chromosome with global cluster indices: out[2*i], out[2*i+1] */
pmmc_status pmmc_tree_arcs(const pmmc_tree *t, int32_t chr, int32_t *out, int32_t cap,
                           int32_t *n);
/* current_level[chr] after setLevel(level), level 0 = the root */
pmmc_status pmmc_tree_level(const pmmc_tree *t, int32_t chr, int32_t level, int32_t *out,
                            int32_t cap, int32_t *n);
pmmc_status pmmc_tree_set_positions(pmmc_tree *t, int32_t n, const int32_t *clusters,
                                    const float *xyz);
pmmc_status pmmc_tree_get_positions(const pmmc_tree *t, int32_t n, const int32_t *clusters,
                                    float *xyz);
pmmc_status pmmc_tree_set_fixed(pmmc_tree *t, int32_t n, const int32_t *clusters,
                                const int32_t *fixed);

/* densifyActiveRegion over the children of `cluster`: inserts loop_density
 * beads between consecutive children (genomic spacing range/(add+1), linear
 * interpolation of positions), appends them to the tree, fixes the children
 * when fix != 0 and makes the densified list the cluster's children.
 * *n_active (nullable) = the new number of children. */
pmmc_status pmmc_tree_densify(pmmc_tree *t, int32_t cluster, int32_t loop_density, int32_t fix,
                              int32_t *n_active);

/* Sets dist_to_next of the children of `cluster` from their genomic spacing
 * and returns the mean spacing times noise_anchor / noise_subanchor (the
 * reference's noise_size; either pointer nullable). */
pmmc_status pmmc_tree_set_dist_to_next(pmmc_tree *t, int32_t cluster, const pmmc_level_params *p,
                                       float *noise_anchor, float *noise_subanchor);

/* Expected distances between the children of interaction block `cluster` of
 * chromosome `chr`: -1 = no arc, 0 on the diagonal, freq_to_distance of the
 * arc's count for arcs (and long arcs resolved to anchors) between children;
 * an arc whose other end lies outside the block is written against child 0,
 * as the reference does.  anchor_heatmap (nullable, n*n) applies the
 * useAnchorHeatmap shrink when p->use_anchor_heatmap.  out: n*n row-major,
 * n = number of children. */
pmmc_status pmmc_tree_expected_distances(const pmmc_tree *t, int32_t chr, int32_t cluster,
                                         const pmmc_level_params *p,
                                         const float *anchor_heatmap, float *out);

/* ---- level drivers ------------------------------------------------------- */
/*
 * The reference runs every restart of every interaction block as its own
 * host loop iteration with its own GPU launch (LooperSolver.hpp:1334-1398,
 * :4027-4138, :4408-4463).  Here all restarts of all blocks of a stage are
 * one batched launch (pmmc_anneal_batch): restart k of block b is job
 * (b, k), whose chains start from the block's initial structure plus that
 * restart's noise; the best restart of each block (lowest score, first on
 * ties, as `score < best_score || best_score < 0`) is written back into the
 * tree.  Random streams: restart noise for bead i of job j is
 * (2u-1)*amplitude per axis from Philox(seed, {i, 0, segment + j, 5<<28})
 * (z = 0 under use2d); the annealing chains of job j run with segment
 * `segment + j`, chain ids from 0.
 */

typedef struct {
  int32_t n_blocks;          /* blocks (or regions) annealed                   */
  int32_t n_jobs;            /* jobs launched over all stages                  */
  int64_t moves;             /* counted moves over all chains                  */
  int64_t accepted;
  double score_anchor;       /* sum over blocks of the best restart's score:   */
  double score_smooth;       /*   anchor / sub-anchor stage (heatmap: score_anchor) */
  double ms_anchor;          /* host wall time of each batched stage          */
  double ms_smooth;
  double score_global;       /* global refinement: sum over chromosomes of the */
  double ms_global;          /*   best restart's score; wall time incl. stress */
  int32_t stress_iterations; /* stress updates summed over chromosomes         */
  int32_t reinterpolated;    /* sub-anchor beads re-placed between anchors     */
} pmmc_level_report;

typedef struct {
  pmmc_level_settings arcs;   /* [simulation_arcs] + [cuda] settings           */
  pmmc_level_settings smooth; /* [simulation_arcs_smooth] settings             */
  int32_t arcs_preset;        /* pmmc_preset of the anchor stage               */
  int32_t smooth_preset;      /* of the sub-anchor stage; blocks of fewer than
                                 4 beads always take the CPU annealer, as the
                                 reference dispatches (LooperSolver.hpp:4431-4443) */
  float k_stretch_arcs, k_squeeze_arcs;   /* springConstantStretchArcs / SqueezeArcs */
  float k_stretch, k_squeeze, k_angular;  /* [springs]                       */
  float w_dist, w_angle;                  /* weightDistSmooth / weightAngleSmooth */
  int32_t chains_per_restart; /* chains of each job; 0 = share the library
                                 default chain pool among the jobs             */
  int32_t densify;            /* 0: stop after the anchor stage                */
  int32_t global_refine;      /* useGlobalAnchorRefinement (default 0)         */
  pmmc_stress_opts stress;    /* stressRefineIterations / LR of the global pass */
  int32_t one_call_per_job;   /* 1: each block restart as its own plain call, in
                                 order, as the reference dispatches them (same
                                 results; for comparison); 0: one batch per stage */
} pmmc_arcs_level_opts;

void pmmc_arcs_level_opts_default(pmmc_arcs_level_opts *o);

/* Arc and sub-anchor levels over the interaction blocks of chromosome `chr`
 * (-1: every chromosome): reconstructClustersArcsDistances' block loop
 * (LooperSolver.hpp:3978-4140) over reconstructClusterArcsDistances
 * (:4303-4487).  Per block: children start at the block's position, expected
 * distances as pmmc_tree_expected_distances, restarts of the arc annealer
 * with step = mean dist_to_next * noise_anchor and restart noise
 * noise_anchor_small on every bead; then densification (fixing the anchors)
 * and restarts of the smooth annealer with step = restart noise = mean
 * dist_to_next * noise_subanchor on the non-fixed beads.  The interaction
 * blocks' own positions are the caller's (positionInteractionBlocks).
 *
 * global_refine (:4141-4243), per chromosome after its blocks: the active
 * region becomes all its anchors (tree order); expected distances over them
 * (arcs and long arcs); restarts of the arc annealer from the current
 * positions (step = mean dist_to_next * noise_anchor, restart noise
 * noise_anchor_small on every bead; anchors the sub-anchor stage fixed stay
 * fixed, so their chains stop at once, PMMC_STOP_NOFREE); then, when
 * stress.max_iter > 0, pmmc_tree_stress_refine's descent over the anchors
 * sorted by genomic position; then every sub-anchor bead (level 5) that
 * existed when the reference's pass ran (the beads of this and earlier
 * chromosomes) whose genomic position lies within the anchors' range is
 * re-placed on the line between the two anchors around it -- whichever
 * chromosome the bead belongs to, as the reference's range test does.
 * Random streams: the global restarts take the segments that follow the
 * chromosome's block calls. */
pmmc_status pmmc_arcs_level(pmmc_ctx *ctx, pmmc_tree *t, int32_t chr,
                            const pmmc_level_params *p, const pmmc_arcs_level_opts *o,
                            uint64_t seed, uint32_t segment, pmmc_level_report *report);

/* LooperSolver::stressRefineAnchors (LooperSolver.hpp:5638-5915) over the
 * anchors of chromosome `chr` (-1: every chromosome, one launch), the active
 * region the global refinement sets: anchors sorted by genomic position;
 * pairs from the chromosome's arcs (targets freq_to_distance of the count),
 * then its long arcs resolved to the first anchor containing each end, the
 * first pair of two anchors kept, none of anchors adjacent in the sorted
 * order while a spring constant exceeds 1e-6, none with a target <= 1e-6;
 * spring targets genomic_length_to_distance of the gap to the next anchor
 * (none for a zero gap); then pmmc_stress_refine, positions written back.
 * stats (nullable) receives one record per chromosome processed. */
pmmc_status pmmc_tree_stress_refine(pmmc_ctx *ctx, pmmc_tree *t, int32_t chr,
                                    const pmmc_level_params *p, float k_stretch, float k_squeeze,
                                    float k_angular, const pmmc_stress_opts *opts,
                                    pmmc_stress_stats *stats);

typedef struct {
  pmmc_level_settings heatmap;  /* [simulation_heatmap] + [cuda] settings       */
  int32_t preset;               /* pmmc_preset                                  */
  int32_t chains_per_restart;   /* 0 = share the library default chain pool     */
  int32_t one_call_per_job;     /* as in pmmc_arcs_level_opts                   */
} pmmc_heatmap_level_opts;

void pmmc_heatmap_level_opts_default(pmmc_heatmap_level_opts *o);

/* Heatmap level over tree level `level` (0: chromosome roots, 1: segments):
 * reconstructClustersHeatmapSingleLevel (LooperSolver.hpp:1258-1422).  The
 * active region is the concatenation of every chromosome's level list; dist
 * is its n x n distance heatmap (pmmc_distance_heatmap) with band
 * `diagonal`.  Noise amplitude = annealing step = getAvg(dist) * noise_chr
 * (level 0) or * noise_segment (level 1, where every segment first moves to
 * its chromosome's position and chromosome boundaries split the heatmap when
 * there are several chromosomes, :1221-1229).  Restarts restarts_chr /
 * restarts_segment; the best restart's structure is written back. */
pmmc_status pmmc_heatmap_level(pmmc_ctx *ctx, pmmc_tree *t, int32_t level, const float *dist,
                               int32_t diagonal, const pmmc_level_params *p,
                               const pmmc_heatmap_level_opts *o, uint64_t seed, uint32_t segment,
                               pmmc_level_report *report);

/* ---- reconstruction-quality metrics (host) -------------------------------- */
/* xyz arrays are [3*n] float, bead-major; maps are [n*n] float row-major. */

/* Pearson correlation with FP64 sums in input order; 0 for n <= 0 or a
 * degenerate denominator <- MetricsFramework::pearsonCorrelation
 * (MetricsFramework.hpp:238-266), BenchmarkRunner::calcDistanceCorrelation
 * (BenchmarkRunner.hpp:340-368). */
float pmmc_pearson(const float *a, const float *b, int64_t n);

/* Pearson correlation of the pairwise bead distances of two structures over
 * the pairs (i, j), i < j, both multiples of `step` <-
 * MetricsFramework::distanceCorrelation (MetricsFramework.hpp:268-286).
 * step <= 0 selects BenchmarkRunner::runSingle's rule, n / 500 when n > 500
 * and 1 otherwise (BenchmarkRunner.hpp:314-331). */
float pmmc_distance_correlation(const float *xyz_a, const float *xyz_b, int32_t n, int32_t step);

/* FP32 pairwise distance map of a structure (0 on the diagonal). */
pmmc_status pmmc_distance_map(const float *xyz, int32_t n, float *out);

/* Stratum-adjusted correlation of two maps: per-diagonal Pearson for
 * k = 1..n-2, weighted by (n-k)·exp(-k²/(2·h²)) <-
 * MetricsFramework::structuralSimilarity / perDiagonalCorrelation
 * (MetricsFramework.hpp:419-479; the reference's default h is 5). */
float pmmc_scc(const float *a, const float *b, int32_t n, int32_t h_param);

/* Least RMSD between two structures over proper rotations and translations
 * <- Chromosome::calcRMSD (Chromosome.hpp:613-634), fast_rmsd
 * (thirdparty/rmsd.cpp:395). */
float pmmc_rmsd(const float *xyz_a, const float *xyz_b, int32_t n);

#ifdef __cplusplus
}
#endif
#endif /* PMMC_LEVELS_H */
