/*
 * pmmc_formats.h -- C ABI of the on-disk formats around the cluster tree:
 * the HCM model files, PDB and mmCIF structure files and .heat heatmaps.
 *
 * Pure C, host code only (nothing here touches the GPU), caller-owned memory,
 * pmmc_status returns as in pmmc_hip.h; the message of the last failure is
 * pmmc_host_last_error() (include/pmmc_levels.h).  Text is written with the C
 * library's printf conversions, so files are byte-identical to the
 * reference's for the same model.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference's src/ directory):
 *   pmmc_tree_write_hcm     HierarchicalChromosome::toFile (new format,
 *                           HierarchicalChromosome.cu:522-552) and
 *                           toFilePreviousFormat (:554-593), over
 *                           Cluster::toFile / toFilePreviousFormat
 *                           (include/Cluster.hpp:175-189) and
 *                           InteractionArc::toFile (include/InteractionArc.hpp:122-125);
 *                           the model is LooperSolver::getModel (clusters, arcs,
 *                           chromosomes and roots, include/LooperSolver.hpp:1997-2005)
 *   pmmc_tree_read_hcm      HierarchicalChromosome::fromFile (format detection
 *                           :595-617, new format :619-674, previous format
 *                           :676-740), Cluster::fromFile (Cluster.hpp:191-210),
 *                           InteractionArc::fromFile (InteractionArc.hpp:127-130)
 *   pmmc_tree_write_pdb     PdbWriter::write (include/PdbWriter.hpp:90-294)
 *   pmmc_write_pdb_flat     PdbWriter::writeFlat (:296-335)
 *   pmmc_tree_write_cif     CifWriter::write (include/CifWriter.hpp:126-298) with
 *                           its helpers (:66-124)
 *   pmmc_write_cif_flat     CifWriter::writeFlat (:300-333)
 *   pmmc_heat_write         Heatmap::toFile (include/Heatmap.hpp:679-700)
 *   pmmc_heat_read          Heatmap::fromFile (:702-732)
 */
#ifndef PMMC_FORMATS_H
#define PMMC_FORMATS_H

#include <stdint.h>

#include "pmmc_levels.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Model fields the files carry beyond the tree's own: the factor names
 * (InteractionArcs::factors, in order) ... */
pmmc_status pmmc_tree_set_factors(pmmc_tree *t, int32_t n, const char *const *names);
/* ... each arc's genomic ends, effective score and factor (as
 * pmmc_group_arcs emits them; start / end / score must be the tree's own,
 * n = the chromosome's arc count).  Without it an arc carries the values of
 * InteractionArc(start, end, score): genomic ends -1, eff_score = score,
 * factor -1 ... */
pmmc_status pmmc_tree_set_arc_info(pmmc_tree *t, int32_t chr, int32_t n, const pmmc_arc *arcs);
/* ... and each cluster's annotation type (Cluster::annotation_type: 0 none,
 * 1 CTCF, 2 enhancer, 3 promoter, 4 epigenetic; default 0), which the PDB
 * writer puts in the B-factor column. */
pmmc_status pmmc_tree_set_annotation(pmmc_tree *t, int32_t n, const int32_t *clusters,
                                     const int32_t *types);

typedef enum {
  PMMC_HCM_PREVIOUS = 0,  /* toFilePreviousFormat: the first chromosome only; the
                             format `create` writes unless asked for the new one */
  PMMC_HCM_NEW = 1,       /* toFile: every chromosome with its root and arcs */
  PMMC_HCM_DETECT = 2     /* reading only: the reference's rule -- more than 100
                             clusters and fewer than 5 in the header's second
                             field read as the new format, else the previous one */
} pmmc_hcm_format;

pmmc_status pmmc_tree_write_hcm(const pmmc_tree *t, const char *path, int32_t format);
/* The tree of an HCM file: clusters (start, end, orientation, position,
 * children; genomic_pos = (start + end) / 2), chromosomes with their roots and
 * arcs, factors; cluster arc lists as the reader builds them.  Levels,
 * parents and anchors are not in the file (0, -1, none). */
pmmc_status pmmc_tree_read_hcm(const char *path, int32_t format, pmmc_tree **out);

/* Structure files of the tree's beads at one level: level < 0 = the lowest
 * level (every chromosome's leaves), else the chromosome roots taken `level`
 * times to their children; each step keeps childless clusters and sorts the
 * chromosome's list by genomic position (HierarchicalChromosome::setLevel /
 * useLowestLevel, HierarchicalChromosome.cu:211-257).  Coordinates are
 * centred and scaled so the farthest bead lies at 50. */
pmmc_status pmmc_tree_write_pdb(const pmmc_tree *t, int32_t level, const char *path);
pmmc_status pmmc_tree_write_cif(const pmmc_tree *t, int32_t level, const char *data_name,
                                const char *path);
/* One chain of positions as given (no scaling). */
pmmc_status pmmc_write_pdb_flat(const float *xyz, int32_t n, const char *chr, const char *path);
pmmc_status pmmc_write_cif_flat(const float *xyz, const int32_t *genomic_pos, int32_t n_pos,
                                int32_t n, const char *chr, const char *path);

/* .heat text heatmap: optional "n" header, n rows of n values ("%f " or
 * "%d " of the truncated value), the diagonal optionally written as 0. */
pmmc_status pmmc_heat_write(const float *h, int32_t n, const char *path, int32_t header,
                            int32_t zero_diagonal, int32_t as_integers);
/* Reads a .heat file whose first line holds 1 field (n) or 3 (n start
 * resolution; the field count is one more than the blanks and tabs of the
 * line's first 63 characters): *n; values [n*n] into out when cap >= n*n
 * (call with out = NULL to size); *diagonal_size (nullable) = the heatmap's
 * diagonal band (Heatmap::getDiagonalSize).  *start / *resolution (nullable)
 * receive what the reference's heatmap holds after reading, -1 and 0: its
 * Heatmap::init, called after the header is parsed, resets both. */
pmmc_status pmmc_heat_read(const char *path, float *out, int64_t cap, int32_t *n, int32_t *start,
                           int32_t *resolution, int32_t *diagonal_size);

#ifdef __cplusplus
}
#endif
#endif
