// Internal launchers shared between the op translation units.
#pragma once

#include "clk_common.hpp"

namespace clk {

// out[row] (+)= sum over col of in[col * rows + row], fixed order.
int launch_reduce_f64(clk_context *ctx, const double *in, uint64_t cols, uint64_t rows,
                      double *out, bool accumulate, hipStream_t stream);
int launch_reduce_u32(clk_context *ctx, const uint32_t *in, uint64_t cols, uint64_t rows,
                      uint32_t *out, bool accumulate, hipStream_t stream);

// sums[r] += sum over col of ps[col*rows_s + r] (FP64) and
// masses[r] += sum over col of pm[col*rows_m + r] (uint32), one launch.
int launch_reduce_partials(clk_context *ctx, const double *ps, uint64_t rows_s, double *sums,
                           const uint32_t *pm, uint64_t rows_m, uint32_t *masses, uint64_t cols,
                           hipStream_t stream);

struct FinalizeArgs;
// Merge per-work-group partials into sums/masses (mode 1 accumulate, mode 2
// overwrite + FP64 divide + optional shift): one work-group launch.
int launch_finalize(clk_context *ctx, const FinalizeArgs &a, hipStream_t stream);

// Smallest register bucket >= d among 1,2,4,8,16,32,64; 0 when d > 64.
inline int dmax_bucket(uint64_t d)
{
    int b = 1;
    while (b < 64 && (uint64_t)b < d)
        b *= 2;
    return ((uint64_t)b >= d) ? b : 0;
}

// Labels every point with the reference arithmetic; used by the fused ops
// for the global-accumulation variant and by clk_labeling.
// bounded: the "unroll_vector_bounded" variant (labels read first, kept
// labels not rescanned; the same labels).
int launch_labeling(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                    uint64_t k, const float *points, const float *centroids, uint32_t *labels,
                    hipStream_t stream, bool bounded = false);

// Labels from x·cᵀ on MFMA (16-bit operands, FP32 accumulation) with an exact
// FP32 relabel of near ties: the same labels as launch_labeling.  Supported
// for d <= 128 and n < 2^32 when the 16-bit centroid table fits in LDS.
bool mfma_labeling_supported(const clk_context *ctx, uint64_t d, uint64_t n, uint64_t k);
// Once per context: whether the matrix cores keep FP16 (bit 0) / BF16 (bit 1)
// subnormal inputs, into ctx->mfma_subnormals (synchronizes the stream).
int mfma_subnormal_flags(clk_context *ctx, hipStream_t stream);
int launch_labeling_mfma(clk_context *ctx, bool fp16, uint64_t d, uint64_t n, uint64_t k,
                         const float *points, const float *centroids, uint32_t *labels,
                         hipStream_t stream);

// The mass and centroid-update stages (accumulate into masses / sums); a
// null cfg selects part_private / cluster_merge.
int run_mass_update(clk_context *ctx, const clk_kernel_config *cfg, uint64_t n, uint64_t k,
                    const uint32_t *labels, uint32_t *masses, hipStream_t stream);
int run_centroid_update(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                        uint64_t k, const float *points, double *sums, const uint32_t *labels,
                        hipStream_t stream);

// sums[f,c] += x_f (the grid value of x_f, clk_grid.hpp) and masses[c] += 1
// for every point, with FP64 / u32 global atomics (any k*d; exact, so the
// order of the atomics does not matter).  masses may be null.
struct GridRef;
int launch_accumulate_global(clk_context *ctx, uint64_t d, uint64_t n, uint64_t k,
                             const float *points, const uint32_t *labels, const GridRef &grid,
                             double *sums, uint32_t *masses, hipStream_t stream);

// One Lloyd iteration through the fused pass (clk_lloyd_step) with a given
// sum grid instead of the pinned or per-call one.  CLK_ERR_UNSUPPORTED, with
// nothing enqueued, when k*d needs the separate stages.
int lloyd_step_with_grid(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n, uint64_t k,
                         const float *points, const float *old_centroids, float *new_centroids, double *sums,
                         uint32_t *labels, uint32_t *masses, double *max_shift, const GridRef &grid,
                         hipStream_t stream);

// Device address of this device's static scratch of clk_lloyd_step_noalloc
// into ctx->noalloc_scratch (hipGetSymbolAddress once per device).
int noalloc_setup(clk_context *ctx);

}  // namespace clk
