// Driver that runs the reference's own CPU InteractionArcs::markArcs on arcs
// read from stdin and prints the anchor-indexed arcs.  Test infrastructure:
// built by oracle/Makefile into oracle/_ref/ only when the reference sources
// are present; the reference's headers are used in place, nothing is copied.
//
// stdin:  C  (chromosomes), then per chromosome: A R, A lines "start end",
//         R lines "start end score factor"
// stdout: per chromosome: K, then K lines
//         "start end genomic_start genomic_end score eff_score factor"
#include <InteractionArcs.hpp>
#include <Static.hpp>

int main() {
  InteractionArcs ia;
  int C;
  if (scanf("%d", &C) != 1) return 1;
  for (int c = 0; c < C; ++c) {
    std::string chr = "chr" + std::to_string(c + 1);
    int A, R;
    if (scanf("%d %d", &A, &R) != 2) return 1;
    ia.chrs.push_back(chr);
    std::vector<Anchor> anchors;
    for (int i = 0; i < A; ++i) {
      int s, e;
      if (scanf("%d %d", &s, &e) != 2) return 1;
      anchors.push_back(Anchor(chr, s, e, 'N'));
    }
    std::vector<InteractionArc> raw;
    for (int i = 0; i < R; ++i) {
      int s, e, sc, f;
      if (scanf("%d %d %d %d", &s, &e, &sc, &f) != 4) return 1;
      raw.push_back(InteractionArc(s, e, sc, f));
    }
    ia.anchors[chr] = anchors;
    ia.anchors_cnt[chr] = A;
    ia.raw_arcs[chr] = raw;
  }
  // markArcs prints progress on stdout; send it to stderr.
  fflush(stdout);
  int saved = dup(1);
  dup2(2, 1);
  ia.markArcs(false);
  fflush(stdout);
  dup2(saved, 1);
  for (int c = 0; c < C; ++c) {
    std::string chr = ia.chrs[c];
    const auto &v = ia.arcs[chr];
    printf("%d\n", (int)v.size());
    for (const auto &a : v)
      printf("%d %d %d %d %d %d %d\n", a.start, a.end, a.genomic_start, a.genomic_end, a.score,
             a.eff_score, a.factor);
  }
  return 0;
}
