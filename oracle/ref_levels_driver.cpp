// Test infrastructure only: runs the reference's own host code, compiled in
// place from its sources, on inputs read from stdin and prints the results as
// text (floats as C99 hex literals, so nothing is lost).  Used to write the
// golden fixtures of tests/golden/ for the host-side rows of the annealing
// path and to cross-check the oracle's energies.  Nothing here is shipped.
//
// Modes (first token on stdin):
//   tree      LooperSolver::createTreeGenome (LooperSolver.hpp:2272-2438)
//             with findGaps / findSplit (:2119-2269), then setLevel(0..4)
//             (:2441-2452, levelDown :3087-3104)
//   densify   tree, then LooperSolver::densifyActiveRegion (:3836-3898) on one
//             interaction block and genomicLengthToDistance (:3900-3904)
//   expdist   tree, then LooperSolver::calcAnchorExpectedDistancesHeatmap
//             (:5494-5627) on one interaction block
//   distheat  LooperSolver::createDistanceHeatmap (:3042-3085)
//   score     calcScoreHeatmapActiveRegion (:3583-3612),
//             calcScoreDistancesActiveRegion (:3215-3247),
//             calcScoreStructureSmooth (:3414-3451)
//   anneal / time_anneal
//             the reference's CPU annealers: MonteCarloHeatmap (:1424-1565),
//             MonteCarloArcs (:4670-4796), MonteCarloArcsSmooth (:4798-5047),
//             driven by glibc rand() after srand(seed), as the reference's
//             `-j seed` option seeds them (main.cpp:2833-2845)
//   metrics   MetricsFramework::pearsonCorrelation, distanceCorrelation,
//             structuralSimilarity (MetricsFramework.hpp:238-479) and
//             Chromosome::calcRMSD -> fast_rmsd (Chromosome.hpp:613-634)
//   singleton tree, then LooperSolver::createSingletonHeatmap (:2517-2749)
//             on records written to a temporary singleton file
//   singleton_clusters  the same on clusters given directly, timed
//   normalize normalizeHeatmap (:2996-3040), normalizeHeatmapDiagonalTotal
//             (:3146-3172), normalizeHeatmapInter (:3106-3144)
//   drive_arcs
//             tree, then the per-block body of reconstructClustersArcsDistances
//             (:4027-4138) for one chromosome: children at the block's
//             position, calcAnchorExpectedDistancesHeatmap,
//             reconstructClusterArcsDistances(ib, i, false) (:4303-4487),
//             densifyActiveRegion, reconstructClusterArcsDistances(ib, i, true);
//             the annealing calls go through the product's reference-side
//             glue to libpmmc_hip.so (a GPU is needed)
//             With a trailing "global ..." record: long arcs, and after each
//             chromosome's blocks the global anchor refinement of
//             :4141-4243 -- calcAnchorExpectedDistancesHeatmap and
//             reconstructClusterArcsDistances(root, 0, false, false) over all
//             anchors, stressRefineAnchors (:5638-5915), then the re-placement
//             of sub-anchor beads between the anchors, which the reference
//             writes inline (:4179-4239) and this driver restates.  The
//             chromosome name "*" drives every chromosome in order.
//   records   Cluster::toFile / toFilePreviousFormat (Cluster.hpp:175-189) and
//             InteractionArc::toFile (InteractionArc.hpp:122-125) of given
//             clusters and arcs, as the HCM writers emit them
//   heat      Heatmap::toFile (Heatmap.hpp:679-700) of a given matrix to a
//             given path, then Heatmap::fromFile (:702-732) of a given path
//   stress    tree, then LooperSolver::stressRefineAnchors (:5638-5915) over
//             one chromosome's anchors with given positions and long arcs
//   flat      PdbWriter::writeFlat (PdbWriter.hpp:296-335) and
//             CifWriter::writeFlat (CifWriter.hpp:300-333, with its header,
//             genomic-metadata and atom-site helpers :66-124) of given
//             positions to given paths.  The writers' model forms (write)
//             need a HierarchicalChromosome, whose members live in a CUDA
//             source file, so they are neither called nor emitted here.
//   drive_heatmap
//             tree, then reconstructClustersHeatmapSingleLevel(level)
//             (:1258-1422) on a given distance heatmap, with the chromosome
//             boundaries its caller sets (:1219-1229), through the glue
//
// The reference keeps these members private; this driver is a white-box test,
// so it lifts the access specifiers before including the reference header.
#include <bits/stdc++.h>
#define private public
#define protected public
#include <LooperSolver.hpp>
#include <MetricsFramework.hpp>
#include <AppCommon.hpp>  // defines g_cancel_requested (read by createSingletonHeatmap)
#undef private
#undef protected
#include <Static.hpp>

#include <PdbWriter.hpp>
#include <CifWriter.hpp>

#include "../pmmc_amd/integration/reference_glue.h"

static bool maybe_tok(std::string &s) { return static_cast<bool>(std::cin >> s); }

static std::string tok() {
  std::string s;
  if (!(std::cin >> s)) {
    fprintf(stderr, "unexpected end of input\n");
    exit(2);
  }
  return s;
}
static int geti() { return std::stoi(tok()); }
static float getf() { return std::strtof(tok().c_str(), nullptr); }

// Input: n_chr, per chromosome "name n_anchors n_arcs", anchors "start end
// orient", arcs "a b score" (anchor indices local to the chromosome), then
// n_seg and "chr start end" predefined segments.
static void read_tree(LooperSolver &s) {
  int n_chr = geti();
  std::vector<std::string> names;
  for (int c = 0; c < n_chr; ++c) {
    std::string chr = tok();
    int na = geti(), nr = geti();
    names.push_back(chr);
    std::vector<Anchor> anc;
    for (int i = 0; i < na; ++i) {
      int st = geti(), en = geti();
      char o = tok()[0];
      anc.push_back(Anchor(chr, st, en, o));
    }
    std::vector<InteractionArc> arcs;
    for (int i = 0; i < nr; ++i) {
      int a = geti(), b = geti(), sc = geti();
      arcs.push_back(InteractionArc(a, b, sc));
    }
    s.arcs.anchors[chr] = anc;
    s.arcs.anchors_cnt[chr] = na;
    s.arcs.arcs[chr] = arcs;
    s.arcs.arcs_cnt[chr] = nr;
  }
  int n_seg = geti();
  for (int i = 0; i < n_seg; ++i) {
    std::string chr = tok();
    int st = geti(), en = geti();
    s.segments_predefined.regions.push_back(BedRegion(chr, st, en));
  }
  s.arcs.chrs = names;
  s.chrs = names;
  s.createTreeGenome();
}

static void print_tree(LooperSolver &s) {
  printf("\nclusters %zu\n", s.clusters.size());
  for (size_t i = 0; i < s.clusters.size(); ++i) {
    Cluster &c = s.clusters[i];
    printf("%d %d %d %d %d %d %d %zu", c.start, c.end, c.genomic_pos, c.base_start, c.base_end,
           c.level, c.parent, c.children.size());
    for (int ch : c.children) printf(" %d", ch);
    printf("\n");
  }
  for (const std::string &chr : s.chrs) {
    printf("chr %s %d %d\n", chr.c_str(), s.chr_first_cluster[chr], s.chr_root[chr]);
    printf("arcs %zu", s.arcs.arcs[chr].size());
    for (auto &a : s.arcs.arcs[chr]) printf(" %d %d", a.start, a.end);
    printf("\n");
  }
  for (int lvl = 0; lvl <= 4; ++lvl) {
    s.setLevel(lvl);
    for (const std::string &chr : s.chrs) {
      printf("level %d %s %zu", lvl, chr.c_str(), s.current_level[chr].size());
      for (int c : s.current_level[chr]) printf(" %d", c);
      printf("\n");
    }
  }
}

static void mode_densify(LooperSolver &s) {
  int ib = geti();
  Settings::loopDensity = geti();
  bool fix = geti() != 0;
  s.active_region = s.clusters[ib].children;
  for (int c : s.active_region) {
    float x = getf(), y = getf(), z = getf();
    s.clusters[c].pos.set(x, y, z);
  }
  s.densifyActiveRegion(ib, fix);
  printf("\nactive %zu\n", s.active_region.size());
  for (int c : s.active_region) {
    Cluster &k = s.clusters[c];
    printf("%d %d %d %d %d %d %a %a %a\n", c, k.start, k.end, k.genomic_pos, k.level,
           (int)k.is_fixed, k.pos.x, k.pos.y, k.pos.z);
  }
  // distances to the next bead as the arc/smooth driver sets them
  // (LooperSolver.hpp:4339-4354)
  float noise = 0.0f;
  printf("dist_to_next");
  for (size_t i = 0; i + 1 < s.active_region.size(); ++i) {
    int d = abs(s.clusters[s.active_region[i + 1]].genomic_pos -
                s.clusters[s.active_region[i]].genomic_pos);
    float dist = static_cast<float>(s.genomicLengthToDistance(d));
    noise += dist;
    printf(" %a", dist);
  }
  printf("\n");
  noise /= s.active_region.size();
  printf("noise %a %a\n", noise * Settings::noiseCoefficientLevelSubanchor,
         noise * Settings::noiseCoefficientLevelAnchor);
  printf("children %zu", s.clusters[ib].children.size());
  for (int c : s.clusters[ib].children) printf(" %d", c);
  printf("\n");
}

static void mode_expdist(LooperSolver &s) {
  std::string chr = tok();
  int ib = geti();
  int n_long = geti();
  for (int i = 0; i < n_long; ++i) {
    int st = geti(), en = geti(), sc = geti();
    s.arcs.long_arcs[chr].push_back(InteractionArc(st, en, sc));
  }
  s.current_chr = chr;
  s.active_region = s.clusters[ib].children;
  int n = (int)s.active_region.size();
  Settings::useAnchorHeatmap = geti() != 0;
  Settings::anchorHeatmapInfluence = getf();
  if (Settings::useAnchorHeatmap) {
    s.heatmap_anchor.init(n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) s.heatmap_anchor.v[i][j] = getf();
  }
  fflush(stdout);
  // the reference prints progress; keep the payload after a marker
  s.calcAnchorExpectedDistancesHeatmap();
  printf("\nexpdist %d\n", n);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) printf("%a ", s.heatmap_exp_dist_anchor.v[i][j]);
    printf("\n");
  }
}

static void mode_distheat(LooperSolver &s) {
  int n = geti(), diag = geti(), ind = geti();
  s.heatmap.resize(std::max(ind + 1, 2));
  s.heatmap[ind].init(n);
  s.heatmap[ind].diagonal_size = diag;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) s.heatmap[ind].v[i][j] = getf();
  s.createDistanceHeatmap(ind);
  printf("\ndistheat %d %d %a\n", n, s.heatmap_dist.diagonal_size, s.heatmap_dist.getAvg());
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) printf("%a ", s.heatmap_dist.v[i][j]);
    printf("\n");
  }
}

// score: n, positions, then heatmap (diag + n*n), then arc expected
// distances (n*n), then dist_to_next (n).
static void mode_score(LooperSolver &s) {
  int n = geti();
  s.clusters.assign(n, Cluster());
  s.active_region.clear();
  for (int i = 0; i < n; ++i) {
    float x = getf(), y = getf(), z = getf();
    s.clusters[i].pos.set(x, y, z);
    s.active_region.push_back(i);
  }
  s.heatmap_dist.init(n);
  s.heatmap_dist.diagonal_size = geti();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) s.heatmap_dist.v[i][j] = getf();
  s.heatmap_exp_dist_anchor.init(n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) s.heatmap_exp_dist_anchor.v[i][j] = getf();
  for (int i = 0; i < n; ++i) s.clusters[i].dist_to_next = getf();
  printf("\nscore %.17g %.17g %.17g\n", s.calcScoreHeatmapActiveRegion(-1),
         s.calcScoreDistancesActiveRegion(), s.calcScoreStructureSmooth(true, true));
  printf("constants %a %a %a %a %a %a %a\n", Settings::springConstantStretchArcs,
         Settings::springConstantSqueezeArcs, Settings::springConstantStretch,
         Settings::springConstantSqueeze, Settings::springAngularConstant,
         Settings::weightDistSmooth, Settings::weightAngleSmooth);
}

// anneal: level seed step_size, then the level's settings
//   (max_temp dt tjs tjc improvement min_successes steps use2d), then
//   heatmap: n diag, positions, fixed flags, n*n heatmap_dist
//   arcs:    n k_stretch k_squeeze, positions, n*n heatmap_exp_dist_anchor
//   smooth:  n k_stretch k_squeeze k_angular w_dist w_angle, positions,
//            fixed flags, dist_to_next
// Prints the returned score, the number of moves and the final positions.
struct AnnealInput {
  int level = 0, n = 0;
  unsigned seed = 0;
  float step = 0.f;
  std::vector<float> pos;
};

// Reads the anneal input (after the mode's own leading fields) into the
// solver state and Settings.
static AnnealInput read_anneal(LooperSolver &s) {
  AnnealInput in;
  in.level = geti();
  const int level = in.level;
  in.seed = static_cast<unsigned>(std::stoul(tok()));
  in.step = getf();
  const float maxT = getf(), dt = getf(), tjs = getf(), tjc = getf(), impr = getf();
  const int min_succ = geti(), steps = geti();
  Settings::use2D = geti() != 0;
  int n = geti();
  in.n = n;
  if (level == 1) {
    Settings::maxTempHeatmap = maxT;
    Settings::dtTempHeatmap = dt;
    Settings::tempJumpScaleHeatmap = tjs;
    Settings::tempJumpCoefHeatmap = tjc;
    Settings::MCstopConditionImprovementHeatmap = impr;
    Settings::MCstopConditionMinSuccessesHeatmap = min_succ;
    Settings::MCstopConditionStepsHeatmap = steps;
  } else if (level == 2) {
    Settings::maxTemp = maxT;
    Settings::dtTemp = dt;
    Settings::tempJumpScale = tjs;
    Settings::tempJumpCoef = tjc;
    Settings::MCstopConditionImprovement = impr;
    Settings::MCstopConditionMinSuccesses = min_succ;
    Settings::MCstopConditionSteps = steps;
  } else {
    Settings::maxTempSmooth = maxT;
    Settings::dtTempSmooth = dt;
    Settings::tempJumpScaleSmooth = tjs;
    Settings::tempJumpCoefSmooth = tjc;
    Settings::MCstopConditionImprovementSmooth = impr;
    Settings::MCstopConditionMinSuccessesSmooth = min_succ;
    Settings::MCstopConditionStepsSmooth = steps;
    Settings::useCTCFMotifOrientation = false;
  }
  int diag = 1;
  if (level == 1) diag = geti();
  if (level == 2) {
    Settings::springConstantStretchArcs = getf();
    Settings::springConstantSqueezeArcs = getf();
  }
  if (level == 3) {
    Settings::springConstantStretch = getf();
    Settings::springConstantSqueeze = getf();
    Settings::springAngularConstant = getf();
    Settings::weightDistSmooth = getf();
    Settings::weightAngleSmooth = getf();
  }
  s.clusters.assign(n, Cluster());
  s.active_region.clear();
  for (int i = 0; i < 3 * n; ++i) in.pos.push_back(getf());
  for (int i = 0; i < n; ++i) {
    s.clusters[i].pos.set(in.pos[3 * i], in.pos[3 * i + 1], in.pos[3 * i + 2]);
    s.active_region.push_back(i);
  }
  if (level != 2)
    for (int i = 0; i < n; ++i) s.clusters[i].is_fixed = geti() != 0;
  if (level == 1) {
    s.heatmap_dist.init(n);
    s.heatmap_dist.diagonal_size = diag;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) s.heatmap_dist.v[i][j] = getf();
    s.heatmap_chromosome_boundaries.clear();
  } else if (level == 2) {
    s.heatmap_exp_dist_anchor.init(n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) s.heatmap_exp_dist_anchor.v[i][j] = getf();
    s.loop_constraints.clear();
  } else {
    for (int i = 0; i < n; ++i) s.clusters[i].dist_to_next = getf();
  }
  return in;
}

static double run_annealer(LooperSolver &s, const AnnealInput &in) {
  if (in.level == 1) return s.MonteCarloHeatmap(in.step);
  if (in.level == 2) return s.MonteCarloArcs(in.step);
  return s.MonteCarloArcsSmooth(in.step, false);
}

// anneal: level seed step_size, then the level's settings
//   (max_temp dt tjs tjc improvement min_successes steps use2d), then
//   heatmap: n diag, positions, fixed flags, n*n heatmap_dist
//   arcs:    n k_stretch k_squeeze, positions, n*n heatmap_exp_dist_anchor
//   smooth:  n k_stretch k_squeeze k_angular w_dist w_angle, positions,
//            fixed flags, dist_to_next
// Prints the returned score, the number of moves and the final positions.
static void mode_anneal(LooperSolver &s) {
  const AnnealInput in = read_anneal(s);
  srand(in.seed);
  const long long before = s.total_mc_steps;
  const double score = run_annealer(s, in);
  printf("\n@@anneal %a %lld\n", score, static_cast<long long>(s.total_mc_steps - before));
  for (int i = 0; i < in.n; ++i) {
    const vector3 &p = s.clusters[i].pos;
    printf("%a %a %a\n", p.x, p.y, p.z);
  }
}

// time_anneal: `seconds`, then the anneal input.  Runs the annealer from the
// same initial positions with seeds seed, seed+1, ... until `seconds` of
// annealing have passed (at least once); prints the runs, the moves and the
// seconds spent inside the annealer (steady_clock), then the score each run
// returned.
static void mode_time_anneal(LooperSolver &s) {
  const double budget = std::stod(tok());
  const AnnealInput in = read_anneal(s);
  long long moves = 0;
  double secs = 0.0;
  int runs = 0;
  std::vector<double> scores;
  do {
    for (int i = 0; i < in.n; ++i)
      s.clusters[i].pos.set(in.pos[3 * i], in.pos[3 * i + 1], in.pos[3 * i + 2]);
    srand(in.seed + runs);
    const long long before = s.total_mc_steps;
    const auto t0 = std::chrono::steady_clock::now();
    scores.push_back(run_annealer(s, in));
    secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    moves += s.total_mc_steps - before;
    ++runs;
  } while (secs < budget);
  printf("\n@@time %d %lld %.9f\n@@scores", runs, moves, secs);
  for (double v : scores) printf(" %a", v);
  printf("\n");
}

// metrics: n h, two structures (n xyz each), two n*n maps, a step for
// distanceCorrelation.
static void mode_metrics() {
  const int n = geti(), h = geti();
  Chromosome a, b;
  for (Chromosome *c : {&a, &b}) {
    for (int i = 0; i < n; ++i) {
      float x = getf(), y = getf(), z = getf();
      c->points.push_back(vector3(x, y, z));
    }
    c->size = n;
  }
  Heatmap ha, hb;
  for (Heatmap *m : {&ha, &hb}) {
    m->init(n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) m->v[i][j] = getf();
  }
  const int step = geti();
  std::vector<float> fa, fb;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      fa.push_back(ha.v[i][j]);
      fb.push_back(hb.v[i][j]);
    }
  MetricsFramework mf;
  const float pear = mf.pearsonCorrelation(fa, fb);
  const float dcor = mf.distanceCorrelation(a, b, step);
  const float scc = mf.structuralSimilarity(ha, hb, h);
  const float rmsd = a.calcRMSD(b);
  printf("\n@@metrics %a %a %a %a\n", pear, dcor, scc, rmsd);
}

// singleton: tree (read_tree), then level diag long_scale long_power
//   region_start region_end, n_centromeres "chr start end", n_long
//   "chr start end score", n_records "chr1 sta stb chr2 enda endb score"
//   (written to a temporary singleton file), then
//   LooperSolver::createSingletonHeatmap (LooperSolver.hpp:2517-2749).
static void singleton_run(LooperSolver &s, int level, int diag, bool use_level);

static void mode_singleton(LooperSolver &s) {
  const int level = geti(), diag = geti();
  singleton_run(s, level, diag, true);
}

// singleton_clusters: n_chr, per chromosome "name n" and n "start end" of
// the level's clusters (no tree), then the singleton input from diag on;
// also prints the seconds spent in createSingletonHeatmap (@@singleton_time).
static void mode_singleton_clusters(LooperSolver &s) {
  const int n_chr = geti();
  s.chrs.clear();
  for (int c = 0; c < n_chr; ++c) {
    const std::string chr = tok();
    const int n = geti();
    s.chrs.push_back(chr);
    std::vector<int> ids;
    for (int i = 0; i < n; ++i) {
      Cluster k;
      k.start = geti();
      k.end = geti();
      ids.push_back(static_cast<int>(s.clusters.size()));
      s.clusters.push_back(k);
    }
    s.current_level[chr] = ids;
  }
  const int diag = geti();
  singleton_run(s, -1, diag, false);
}

static void singleton_run(LooperSolver &s, int level, int diag, bool use_level) {
  Settings::longPETClustersEffectScale = getf();
  Settings::longPETClustersEffectPower = getf();
  const int rs = geti(), re = geti();
  s.selected_region = BedRegion("sel", rs, re);
  s.selected_region.start = rs;
  s.selected_region.end = re;
  const int n_cent = geti();
  for (int i = 0; i < n_cent; ++i) {
    std::string chr = tok();
    int st = geti(), en = geti();
    s.centromeres.regions.push_back(BedRegion(chr, st, en));
  }
  const int n_long = geti();
  for (int i = 0; i < n_long; ++i) {
    std::string chr = tok();
    int st = geti(), en = geti(), sc = geti();
    s.arcs.long_arcs[chr].push_back(InteractionArc(st, en, sc));
  }
  const int n_rec = geti();
  char path[] = "/tmp/pmmc_singletons_XXXXXX";
  const int fd = mkstemp(path);
  FILE *f = fdopen(fd, "w");
  for (int i = 0; i < n_rec; ++i) {
    std::string c1 = tok();
    int sta = geti(), stb = geti();
    std::string c2 = tok();
    int enda = geti(), endb = geti(), sc = geti();
    fprintf(f, "%s %d %d %s %d %d %d\n", c1.c_str(), sta, stb, c2.c_str(), enda, endb, sc);
  }
  fclose(f);
  Settings::dataSplitSingletonFilesByChr = false;
  s.arcs_singletons = {path};
  s.arcs_singletons_inter.clear();
  if (use_level) s.setLevel(level);
  fflush(stdout);
  const auto t0 = std::chrono::steady_clock::now();
  Heatmap h = s.createSingletonHeatmap(diag);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  unlink(path);
  printf("\n@@singleton_time %.6f %d\n", secs, n_rec);
  printf("\n@@singleton %zu\n", h.size);
  for (size_t i = 0; i < h.size; ++i) {
    for (size_t j = 0; j < h.size; ++j) printf("%a ", h.v[i][j]);
    printf("\n");
  }
}

// normalize: n_chr, bins per chromosome, val scale, then the n x n heatmap;
// prints normalizeHeatmap (LooperSolver.hpp:2996-3040), then
// normalizeHeatmapDiagonalTotal (:3146-3172) and normalizeHeatmapInter
// (:3106-3144) each applied to a copy of the input.
static void mode_normalize(LooperSolver &s) {
  const int n_chr = geti();
  int n = 0;
  s.chrs.clear();
  for (int c = 0; c < n_chr; ++c) {
    const std::string chr = "chr" + std::to_string(c + 1);
    const int k = geti();
    s.chrs.push_back(chr);
    s.current_level[chr] = std::vector<int>(k, 0);
    n += k;
  }
  const float val = getf(), scale = getf();
  Heatmap h;
  h.init(n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) h.v[i][j] = getf();
  fflush(stdout);
  Heatmap a = s.normalizeHeatmap(h);
  Heatmap b = h;
  s.normalizeHeatmapDiagonalTotal(b, val);
  Heatmap c = h;
  s.normalizeHeatmapInter(c, scale);
  printf("\n@@normalize %d\n", n);
  for (Heatmap *m : {&a, &b, &c}) {
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) printf("%a ", m->v[i][j]);
      printf("\n");
    }
  }
}

// Annealing settings of one [simulation_*] section: maxT dt tjs tjc
// improvement min_successes steps.
struct AnnealSettings {
  float maxT, dt, tjs, tjc, impr;
  int min_succ, steps;
};

static AnnealSettings read_anneal_settings() {
  AnnealSettings a;
  a.maxT = getf();
  a.dt = getf();
  a.tjs = getf();
  a.tjc = getf();
  a.impr = getf();
  a.min_succ = geti();
  a.steps = geti();
  return a;
}

// harness seed, first Philox segment and chains per call of the glue
static void read_run() {
  const unsigned seed = static_cast<unsigned>(std::stoul(tok()));
  const uint32_t segment = static_cast<uint32_t>(std::stoul(tok()));
  const int chains = geti();
  Settings::gpuSeed = seed;
  setGpuSeed(seed);
  pmmc_glue_set_segment(segment);
  pmmc_glue_set_chains(chains);
  srand(seed);
}

static void print_clusters(LooperSolver &s) {
  printf("clusters %zu\n", s.clusters.size());
  for (size_t i = 0; i < s.clusters.size(); ++i) {
    const Cluster &c = s.clusters[i];
    printf("%a %a %a %d %a %zu", c.pos.x, c.pos.y, c.pos.z, (int)c.is_fixed,
           static_cast<double>(c.dist_to_next), c.children.size());
    for (int ch : c.children) printf(" %d", ch);
    printf("\n");
  }
}

// drive_arcs: tree, chromosome name, then
//   loop_density noise_anchor noise_subanchor noise_anchor_small
//   restarts_anchor restarts_subanchor use2d milestone_fails,
//   [simulation_arcs] and [simulation_arcs_smooth] settings,
//   springs: ks_arcs kq_arcs ks kq ka w_dist w_angle,
//   seed segment chains, then the positions of the chromosome's
//   interaction-block clusters (level list 2, in order)
static void mode_drive_arcs(LooperSolver &s) {
  const std::string chr = tok();
  Settings::loopDensity = geti();
  Settings::noiseCoefficientLevelAnchor = getf();
  Settings::noiseCoefficientLevelSubanchor = getf();
  Settings::noiseAnchorSmall = getf();
  Settings::simulationStepsLevelAnchor = geti();
  Settings::simulationStepsLevelSubanchor = geti();
  Settings::use2D = geti() != 0;
  Settings::milestoneFailsThreshold = geti();
  const AnnealSettings a = read_anneal_settings(), m = read_anneal_settings();
  Settings::maxTemp = a.maxT;
  Settings::dtTemp = a.dt;
  Settings::tempJumpScale = a.tjs;
  Settings::tempJumpCoef = a.tjc;
  Settings::MCstopConditionImprovement = a.impr;
  Settings::MCstopConditionMinSuccesses = a.min_succ;
  Settings::MCstopConditionSteps = a.steps;
  Settings::maxTempSmooth = m.maxT;
  Settings::dtTempSmooth = m.dt;
  Settings::tempJumpScaleSmooth = m.tjs;
  Settings::tempJumpCoefSmooth = m.tjc;
  Settings::MCstopConditionImprovementSmooth = m.impr;
  Settings::MCstopConditionMinSuccessesSmooth = m.min_succ;
  Settings::MCstopConditionStepsSmooth = m.steps;
  Settings::springConstantStretchArcs = getf();
  Settings::springConstantSqueezeArcs = getf();
  Settings::springConstantStretch = getf();
  Settings::springConstantSqueeze = getf();
  Settings::springAngularConstant = getf();
  Settings::weightDistSmooth = getf();
  Settings::weightAngleSmooth = getf();
  Settings::useSubanchorHeatmap = false;
  Settings::useAnchorHeatmap = false;
  Settings::useCTCFMotifOrientation = false;
  read_run();
  std::vector<std::string> chrs;
  if (chr == "*")
    chrs = s.chrs;
  else
    chrs.push_back(chr);
  s.setLevel(LVL_INTERACTION_BLOCK);
  std::vector<int> all_ibs;
  for (const std::string &c : chrs)
    for (int ib : s.current_level[c]) {
      float x = getf(), y = getf(), z = getf();
      s.clusters[ib].pos.set(x, y, z);
      all_ibs.push_back(ib);
    }
  // optional: global <on> <stress iterations> <stress lr> <n long arcs>
  // followed by "chr start end count" per long arc
  std::string g;
  if (maybe_tok(g) && g == "global") {
    Settings::useGlobalAnchorRefinement = geti() != 0;
    Settings::stressRefineIterations = geti();
    Settings::stressRefineLR = getf();
    const int n_long = geti();
    for (int i = 0; i < n_long; ++i) {
      const std::string c = tok();
      const int st = geti(), en = geti(), sc = geti();
      s.arcs.long_arcs[c].push_back(InteractionArc(st, en, sc));
    }
  }
  s.gpu_cache.init();
  fflush(stdout);
  for (const std::string &c : chrs) {
    s.current_chr = c;
    const std::vector<int> ibs = s.current_level[c];
    for (size_t i = 0; i < ibs.size(); ++i) {
      const int ib = ibs[i];
      s.active_region.assign(s.clusters[ib].children.begin(), s.clusters[ib].children.end());
      for (int k : s.active_region) s.clusters[k].pos = s.clusters[ib].pos;
      s.calcAnchorExpectedDistancesHeatmap();
      s.reconstructClusterArcsDistances(ib, static_cast<int>(i), false);
      s.densifyActiveRegion(ib, true);
      s.reconstructClusterArcsDistances(ib, static_cast<int>(i), true, false);
    }
    if (!Settings::useGlobalAnchorRefinement) continue;
    // the global pass, as the reference's chromosome loop runs it after the
    // blocks (LooperSolver.hpp:4148-4178)
    s.active_region.clear();
    const int cstart = s.chr_first_cluster[c];
    const int n_anc = s.arcs.anchors_cnt[c];
    for (int i = 0; i < n_anc; ++i) s.active_region.push_back(cstart + i);
    if (s.active_region.size() <= 1) continue;
    s.calcAnchorExpectedDistancesHeatmap();
    s.reconstructClusterArcsDistances(s.chr_root[c], 0, false, false);
    if (Settings::stressRefineIterations > 0)
      s.stressRefineAnchors(c, Settings::stressRefineIterations, Settings::stressRefineLR);
    // sub-anchor beads (the anchors' level + 1) whose genomic position lies
    // in [first, last] anchor position (active_region's order) move onto
    // the segment between the anchors around them (:4179-4239)
    const int sub_level = s.clusters[s.active_region[0]].level + 1;
    std::vector<int> gp;
    for (int a : s.active_region) gp.push_back(s.clusters[a].genomic_pos);
    for (size_t k = 0; k < s.clusters.size(); ++k) {
      Cluster &b = s.clusters[k];
      if (b.level != sub_level || b.genomic_pos < gp.front() || b.genomic_pos > gp.back()) continue;
      const size_t j = std::lower_bound(gp.begin(), gp.end(), b.genomic_pos) - gp.begin();
      if (j == 0) {
        b.pos = s.clusters[s.active_region.front()].pos;
      } else if (j >= gp.size()) {
        b.pos = s.clusters[s.active_region.back()].pos;
      } else {
        const float u = gp[j] > gp[j - 1]
                            ? static_cast<float>(b.genomic_pos - gp[j - 1]) / static_cast<float>(gp[j] - gp[j - 1])
                            : 0.5f;
        const vector3 &a0 = s.clusters[s.active_region[j - 1]].pos;
        const vector3 &a1 = s.clusters[s.active_region[j]].pos;
        b.pos.x = (1.0f - u) * a0.x + u * a1.x;
        b.pos.y = (1.0f - u) * a0.y + u * a1.y;
        b.pos.z = (1.0f - u) * a0.z + u * a1.z;
      }
    }
  }
  printf("\n@@drive %d %zu\n", s.total_mc_steps, all_ibs.size());
  for (int ib : all_ibs) printf("%d ", ib);
  printf("\n");
  print_clusters(s);
}

// records: n, then per cluster "start end orientation x y z k children...",
// then m, then per arc "start end genomic_start genomic_end score eff_score
// factor".  Prints each record set after a marker line.
static void mode_records() {
  const int n = geti();
  std::vector<Cluster> cl(n);
  for (int i = 0; i < n; ++i) {
    Cluster &c = cl[i];
    c.start = geti();
    c.end = geti();
    c.orientation = tok()[0];
    const float x = getf(), y = getf(), z = getf();
    c.pos.set(x, y, z);
    const int k = geti();
    for (int j = 0; j < k; ++j) c.children.push_back(geti());
  }
  const int m = geti();
  std::vector<InteractionArc> arcs(m);
  for (int i = 0; i < m; ++i) {
    InteractionArc &a = arcs[i];
    a.start = geti();
    a.end = geti();
    a.genomic_start = geti();
    a.genomic_end = geti();
    a.score = geti();
    a.eff_score = geti();
    a.factor = geti();
  }
  fflush(stdout);
  printf("@@cluster\n");
  fflush(stdout);
  for (Cluster &c : cl) c.toFile(stdout);
  fflush(stdout);
  printf("@@cluster_previous\n");
  fflush(stdout);
  for (Cluster &c : cl) c.toFilePreviousFormat(stdout);
  fflush(stdout);
  printf("@@arc\n");
  fflush(stdout);
  for (InteractionArc &a : arcs) a.toFile(stdout);
  fflush(stdout);
  printf("@@end\n");
}

// heat: path_out n total_count zero_diagonal as_integers, n*n values, then
// path_in.  Writes the matrix to path_out with Heatmap::toFile, then reads
// path_in with Heatmap::fromFile and prints size, start, resolution,
// diagonal size and the values.
static void mode_heat() {
  const std::string out = tok();
  const int n = geti();
  const bool total = geti() != 0, zero = geti() != 0, ints = geti() != 0;
  Heatmap h;
  h.init(n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) h.v[i][j] = getf();
  h.toFile(out, total, zero, ints);
  const std::string in = tok();
  Heatmap r;
  r.start = -1;
  r.resolution = -1;
  r.fromFile(in);
  printf("\n@@heat %zu %d %d %d\n", r.size, r.start, r.resolution, r.diagonal_size);
  for (size_t i = 0; i < r.size; ++i) {
    for (size_t j = 0; j < r.size; ++j) printf("%a ", r.v[i][j]);
    printf("\n");
  }
}

// flat: pdb_path cif_path chr n n_pos, 3n position floats, n_pos genomic
// positions
static void mode_flat() {
  const std::string pdb = tok(), cif = tok(), chr = tok();
  const int n = geti(), n_pos = geti();
  std::vector<vector3> pos(n);
  for (int i = 0; i < n; ++i) {
    const float x = getf(), y = getf(), z = getf();
    pos[i] = vector3(x, y, z);
  }
  std::vector<int> gpos(n_pos);
  for (int i = 0; i < n_pos; ++i) gpos[i] = geti();
  if (!PdbWriter::writeFlat(pos, chr, pdb) || !CifWriter::writeFlat(pos, gpos, chr, cif)) exit(3);
}

// stress: tree, chromosome name, springs "ks kq ka", "iterations lr", long
// arcs ("n" then "start end count" each), then the positions of the
// chromosome's anchors in cluster order.  Prints the refined active region
// (sorted by the function) and every anchor's position.
static void mode_stress(LooperSolver &s) {
  const std::string chr = tok();
  Settings::springConstantStretch = getf();
  Settings::springConstantSqueeze = getf();
  Settings::springAngularConstant = getf();
  const int iters = geti();
  const float lr = getf();
  const int n_long = geti();
  for (int i = 0; i < n_long; ++i) {
    const int st = geti(), en = geti(), sc = geti();
    s.arcs.long_arcs[chr].push_back(InteractionArc(st, en, sc));
  }
  const int cstart = s.chr_first_cluster[chr];
  const int n_anc = s.arcs.anchors_cnt[chr];
  s.active_region.clear();
  for (int i = 0; i < n_anc; ++i) {
    float x = getf(), y = getf(), z = getf();
    s.clusters[cstart + i].pos.set(x, y, z);
    s.active_region.push_back(cstart + i);
  }
  s.current_chr = chr;
  fflush(stdout);
  const auto t0 = std::chrono::steady_clock::now();
  s.stressRefineAnchors(chr, iters, lr);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  printf("\n@@stress %zu %.6f\n", s.active_region.size(), ms);
  for (int c : s.active_region) printf("%d ", c);
  printf("\n");
  for (int i = 0; i < n_anc; ++i) {
    const vector3 &p = s.clusters[cstart + i].pos;
    printf("%a %a %a\n", p.x, p.y, p.z);
  }
}

// drive_heatmap: tree, then level (0 or 1), noise_chr noise_segment
//   restarts_chr restarts_segment use2d milestone_fails, [simulation_heatmap]
//   settings, seed segment chains, diag n and the n x n distance heatmap of
//   the level's clusters, then the positions of the chromosome clusters
//   (level list 0, one per chromosome)
static void mode_drive_heatmap(LooperSolver &s) {
  const int level = geti();
  Settings::noiseCoefficientLevelChr = getf();
  Settings::noiseCoefficientLevelSegment = getf();
  Settings::simulationStepsLevelChr = geti();
  Settings::simulationStepsLevelSegment = geti();
  Settings::use2D = geti() != 0;
  Settings::milestoneFailsThreshold = geti();
  const AnnealSettings h = read_anneal_settings();
  Settings::maxTempHeatmap = h.maxT;
  Settings::dtTempHeatmap = h.dt;
  Settings::tempJumpScaleHeatmap = h.tjs;
  Settings::tempJumpCoefHeatmap = h.tjc;
  Settings::MCstopConditionImprovementHeatmap = h.impr;
  Settings::MCstopConditionMinSuccessesHeatmap = h.min_succ;
  Settings::MCstopConditionStepsHeatmap = h.steps;
  Settings::useDensity = false;
  Settings::useTelomerePositions = false;
  Settings::initMethod = "random";
  read_run();
  const int diag = geti(), n = geti();
  s.heatmap_dist.init(n);
  s.heatmap_dist.diagonal_size = diag;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) s.heatmap_dist.v[i][j] = getf();
  s.setLevel(LVL_CHROMOSOME);
  for (const std::string &chr : s.chrs) {
    const int root = s.current_level[chr][0];
    float x = getf(), y = getf(), z = getf();
    s.clusters[root].pos.set(x, y, z);
  }
  s.heatmap_chromosome_boundaries.clear();
  if (level == LVL_SEGMENT && s.chrs.size() > 1) {
    s.setLevel(LVL_SEGMENT);
    int p = 0;
    s.heatmap_chromosome_boundaries.push_back(p);
    for (const std::string &chr : s.chrs) {
      p += static_cast<int>(s.current_level[chr].size());
      s.heatmap_chromosome_boundaries.push_back(p);
    }
  }
  s.gpu_cache.init();
  fflush(stdout);
  s.reconstructClustersHeatmapSingleLevel(level);
  printf("\n@@drive %d %zu\n", s.total_mc_steps, s.active_region.size());
  for (int c : s.active_region) printf("%d ", c);
  printf("\n");
  print_clusters(s);
}

int main() {
  Settings settings;
  settings.init();
  Settings::outputLevel = 0;
  std::string mode = tok();
  // the memo of freqToDistance is filled in the constructor, so the
  // count-to-distance parameters come first (LooperSolver.hpp:950-951)
  Settings::countToDistA = getf();
  Settings::countToDistScale = getf();
  Settings::countToDistShift = getf();
  Settings::countToDistBaseLevel = getf();
  Settings::freqToDistHeatmapScale = getf();
  Settings::freqToDistHeatmapPower = getf();
  Settings::freqToDistHeatmapScaleInter = getf();
  Settings::freqToDistHeatmapPowerInter = getf();
  Settings::heatmapDistanceHeatmapStretching = getf();
  Settings::genomicLengthToDistBase = getf();
  Settings::genomicLengthToDistScale = getf();
  Settings::genomicLengthToDistPower = getf();
  // the constructor insists on a segment-split file (LooperSolver.hpp:977-982);
  // an empty one loads no region, and the regions come from stdin instead
  Settings::dataSegmentsSplit = "/dev/null";
  LooperSolver s("ref", "./");
  if (mode == "tree") {
    read_tree(s);
    print_tree(s);
  } else if (mode == "densify") {
    read_tree(s);
    mode_densify(s);
  } else if (mode == "expdist") {
    read_tree(s);
    mode_expdist(s);
  } else if (mode == "distheat") {
    mode_distheat(s);
  } else if (mode == "score") {
    mode_score(s);
  } else if (mode == "anneal") {
    mode_anneal(s);
  } else if (mode == "time_anneal") {
    mode_time_anneal(s);
  } else if (mode == "metrics") {
    mode_metrics();
  } else if (mode == "singleton") {
    read_tree(s);
    mode_singleton(s);
  } else if (mode == "singleton_clusters") {
    mode_singleton_clusters(s);
  } else if (mode == "normalize") {
    mode_normalize(s);
  } else if (mode == "drive_arcs") {
    read_tree(s);
    mode_drive_arcs(s);
  } else if (mode == "records") {
    mode_records();
  } else if (mode == "heat") {
    mode_heat();
  } else if (mode == "stress") {
    read_tree(s);
    mode_stress(s);
  } else if (mode == "flat") {
    mode_flat();
  } else if (mode == "drive_heatmap") {
    read_tree(s);
    mode_drive_heatmap(s);
  } else {
    fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  }
  return 0;
}
