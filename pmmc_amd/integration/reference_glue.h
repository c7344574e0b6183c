// reference_glue.h -- knobs of the reference-side glue (reference_glue.cpp)
// that a pMMC build or a test harness may set.  The reference's own code calls
// none of them; without them the glue behaves as the reference's .cu members
// (library-default chain count, Philox segments 0, 1, 2, ... per call).
#pragma once

#include <stdint.h>

// chains per annealing call; 0 = the library default (pmmc_default_chains)
void pmmc_glue_set_chains(int32_t chains);
// Philox segment of the next annealing call; each call takes the next one
void pmmc_glue_set_segment(uint32_t segment);
