// Merge of per-work-group partials and, for a whole iteration, the FP64
// divide / centroid-shift step -- one small kernel launched after the
// point-streaming kernel on the same stream.
//
// (A variant that merged inside the streaming kernel, with the last
// work-group to arrive doing the work, was measured slower on gfx950: every
// work-group's agent-scope release fence writes back the XCD's L2, which
// holds the freshly written labels.)
//
// The FP64 addition order is fixed by the grid shape: deterministic.
#pragma once

#include "clk_common.hpp"

namespace clk {

struct FinalizeArgs {
    const double *part_sums;     // [rows][d*k] per-work-group partial sums
    const uint32_t *part_masses; // [rows][k]
    uint32_t rows;               // number of partial rows (work-groups)
    double *sums;                // [d][k] out
    uint32_t *masses;            // [k] out
    float *centroids;            // mode 2: [d][k] out
    const float *old_centroids;  // mode 2: for the shift (may be null)
    double *max_shift;           // mode 2: optional out
    uint32_t d, k;
    uint32_t mode;               // 1 accumulate, 2 overwrite + divide
};

constexpr int kFinalizeThreads = 1024;

}  // namespace clk
