// Host-side tree and helpers shared by levels.cpp (tree, densify, expected
// distances) and drivers.cpp (level drivers).
#pragma once
#include <array>
#include <string>
#include <vector>

#include "pmmc_levels.h"

// Cluster (Cluster.hpp:86-102) as the host rows need it
struct TCluster {
  int32_t start = 0, end = 0, genomic_pos = 0;
  int32_t base_start = 0, base_end = 0;
  int32_t level = 0;
  int32_t parent = -1;
  bool is_fixed = false;
  char orientation = 'N';
  float pos[3] = {0.0f, 0.0f, 0.0f};
  double dist_to_next = 0.0;
  int32_t annotation = 0;  // Cluster::annotation_type
  std::vector<int32_t> children;
  std::vector<int32_t> arcs;  // chromosome-local arc indices touching this cluster
};

struct TChr {
  std::string name;
  int32_t first = 0, root = -1;
  std::vector<std::array<int32_t, 2>> anchors;    // bp
  std::vector<char> orient;
  std::vector<std::array<int32_t, 2>> arcs;       // global cluster indices once built
  std::vector<int32_t> scores;
  // per arc: genomic start, genomic end, effective score, factor
  // (InteractionArc fields the files carry)
  std::vector<std::array<int32_t, 4>> arc_info;
  std::vector<std::array<int32_t, 3>> long_arcs;  // bp, bp, count
};

struct TRegion {
  std::string chr;
  int32_t start, end;
};

struct pmmc_tree {
  std::vector<TCluster> cl;
  std::vector<TChr> chrs;
  std::vector<TRegion> regions;
  std::vector<std::string> factors;  // InteractionArcs::factors
};

namespace pmmc_host {
// first anchor of a chromosome containing a genomic position, -1 for none
struct AnchorLookup {
  explicit AnchorLookup(const TChr &chr);
  int operator()(int32_t x) const;
  const TChr &c;
  bool disjoint = true;
};

extern thread_local std::string g_err;
pmmc_status fail(const std::string &msg, pmmc_status st);
bool valid_cluster(const pmmc_tree *t, int32_t c);
std::vector<int32_t> level_list(const pmmc_tree &t, int32_t chr, int32_t level);
void densify(pmmc_tree &t, int32_t cluster, int32_t add, bool fix);
float set_dist_to_next(pmmc_tree &t, const std::vector<int32_t> &active, const pmmc_level_params *p);
bool expected_distances(const pmmc_tree &t, int32_t chr, const std::vector<int32_t> &active,
                        const pmmc_level_params *p, const float *anchor_heatmap, float *m);
int diagonal_size(const float *h, int n);
float heat_avg(const float *h, int n);
float heat_max(const float *h, int n);
}  // namespace pmmc_host
