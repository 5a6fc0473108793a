// Fused single pass (labeling + mass histogram + centroid feature sums in one
// read of the points): the work of lloyd_fused_cluster_merge.cl and
// lloyd_fused_feature_sum.cl, designed for gfx950.
//
// Layout per work-group (LDS):
//   cen   [k][DMAX]            centroids, zero-padded to the register bucket
//   tile  [k*(d+1)][COLS]      row c*(d+1)+0 = mass of c, rows c*(d+1)+1+f =
//                              sum of feature f of c; COLS = work-group size
//                              (cluster_merge: one column per work-item, so
//                              ds_add never collides) or 1 (feature_sum: one
//                              tile shared by the work-group).
// Each work-item labels VEC consecutive points per step with 128-bit loads
// (one per feature row), writes the labels, and adds the points into its
// column (cluster_merge: plain read-add-write, the column is the work-item's
// own; feature_sum: ds_add_f32 / ds_add_u32 atomics on the shared tile).  At the end the
// tile columns are summed in FP64 in a fixed order and written as per-group
// partials; a fixed-order column reduction adds the partials into the
// caller's FP64 sums and uint32 masses.
#include "clk_common.hpp"
#include "clk_finalize.hpp"
#include "clk_kernels.hpp"

#include <algorithm>

namespace clk {
namespace {

template <int DMAX, int VEC, int BLOCK, bool PRIVATE, bool EXACT, bool BOUNDED>
__global__ void __launch_bounds__(BLOCK > 0 ? BLOCK : (DMAX >= 64 ? 512 : 1024))
fused_tile_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                  const float *__restrict__ cen, uint32_t *__restrict__ labels,
                  double *__restrict__ part_sums, uint32_t *__restrict__ part_masses)
{
    extern __shared__ __align__(16) float smem[];
    const uint32_t B = BLOCK > 0 ? BLOCK : blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t cols = PRIVATE ? B : 1;
    const uint32_t rows = k * (d + 1);
    float *cen_lds = smem;
    float *thr_lds = smem + ((k * DMAX + 3u) & ~3u);  // BOUNDED: keep thresholds [k]
    float *tile = thr_lds + (BOUNDED ? ((k + 3u) & ~3u) : 0u);

    stage_centroids<DMAX>(cen, d, k, cen_lds);
    for (uint32_t i = tid; i < rows * cols; i += B)
        tile[i] = 0.0f;
    __syncthreads();
    if constexpr (BOUNDED) {
        stage_keep_thresholds<DMAX>(cen_lds, k, thr_lds);
        __syncthreads();
    }

    const uint64_t groups = n / VEC;
    const uint64_t stride = (uint64_t)gridDim.x * B;
    float *mine = tile + (PRIVATE ? tid : 0);
    const uint32_t row_stride = (d + 1) * cols;  // < 2^24: the tile fits in LDS
    // Software pipeline, two register buffers: one step's points are in
    // flight while the other step's are labelled and accumulated.
    auto step = [&](const float (&x)[DMAX][VEC], const uint32_t (&prev)[VEC], uint64_t g) {
        const uint64_t p = g * VEC;
        uint32_t lab[VEC];
        if constexpr (BOUNDED) {
            // Keep the previous label where the point is inside its centroid's
            // keep radius (keep_threshold: the full scan provably returns it);
            // otherwise scan every centroid.  Only changed labels are stored.
            bool keep = true;
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                const bool valid = prev[v] < k;
                const uint32_t a = valid ? prev[v] : 0u;
                const float *ca = cen_lds + a * DMAX;
                float t = x[0][v] - ca[0];
                float dist = t * t;
#pragma unroll
                for (int f = 1; f < DMAX; ++f) {
                    t = x[f][v] - ca[f];
                    dist = __builtin_fmaf(t, t, dist);
                }
                keep = keep && valid && dist <= thr_lds[a];
                lab[v] = a;
            }
            if (!keep)
                argmin_lds<DMAX, VEC>(x, cen_lds, k, lab);
            bool changed = false;
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                changed = changed || lab[v] != prev[v];
            if (changed)
                store_labels<VEC>(labels, p, lab);
        } else {
            argmin_lds<DMAX, VEC>(x, cen_lds, k, lab);
            store_labels<VEC>(labels, p, lab);
        }
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            float *t = mine + __umul24(lab[v], row_stride);
            if (PRIVATE) {
                // u32 LDS atomics are full speed (unlike ds_add_f32): one
                // ds_add_u32, no VALU, for the mass counter.
                lds_add(reinterpret_cast<uint32_t *>(t), 1u);
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (EXACT || f < (int)d)
                        lds_acc(t + (f + 1) * cols, x[f][v]);
            } else {
                lds_add(reinterpret_cast<uint32_t *>(t), 1u);
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (EXACT || f < (int)d)
                        lds_add(t + (f + 1) * cols, x[f][v]);
            }
        }
    };
    // Loads are unconditional (index clamped to the last group, always in
    // bounds) so the loop body is straight-line and the wait counters can
    // keep the prefetch in flight across the labeling of the other buffer.
    // BOUNDED also prefetches the previous labels of the same points.
    const uint64_t last = groups ? groups - 1 : 0;
    uint64_t g = (uint64_t)blockIdx.x * B + tid;
    float xa[DMAX][VEC], xb[DMAX][VEC];
    uint32_t la[VEC] = {}, lb[VEC] = {};
    auto load = [&](uint64_t gg, float (&x)[DMAX][VEC], uint32_t (&l)[VEC]) {
        load_points<DMAX, VEC, EXACT>(pts, n, d, gg * VEC, x);
        if constexpr (BOUNDED)
            stream_load_labels<VEC>(labels, gg * VEC, l);
    };
    if (groups)
        load(min(g, last), xa, la);
    while (g < groups) {
        const uint64_t g1 = g + stride;
        load(min(g1, last), xb, lb);
        step(xa, la, g);
        if (g1 >= groups)
            break;
        const uint64_t g2 = g1 + stride;
        load(min(g2, last), xa, la);
        step(xb, lb, g1);
        g = g2;
    }
    __syncthreads();

    const uint64_t dk = (uint64_t)d * k;
    for (uint32_t r = tid; r < rows; r += B) {
        const uint32_t c = r / (d + 1), j = r % (d + 1);
        const float *row = tile + (uint64_t)r * cols;
        if (j == 0) {
            part_masses[(uint64_t)blockIdx.x * k + c] =
                PRIVATE ? lds_row_sum_u32(reinterpret_cast<const uint32_t *>(row), cols, r)
                        : reinterpret_cast<const uint32_t *>(row)[0];
        } else {
            part_sums[(uint64_t)blockIdx.x * dk + (uint64_t)(j - 1) * k + c] =
                PRIVATE ? lds_row_sum_f32(row, cols, r) : (double)row[0];
        }
    }
}

using FusedKernel = void (*)(const float *, uint64_t, uint32_t, uint32_t, const float *,
                             uint32_t *, double *, uint32_t *);

// Points per work-item at most (VEC x DMAX values in registers, no spills).
constexpr int fused_vec_cap(int dmax) { return dmax >= 32 ? 1 : dmax >= 16 ? 2 : 4; }

template <int DMAX, int BLOCK, bool PRIVATE, bool EXACT, bool BOUNDED>
FusedKernel pick_vec_block(int vec)
{
    constexpr int cap = fused_vec_cap(DMAX);
    if constexpr (cap >= 4)
        if (vec >= 4)
            return fused_tile_kernel<DMAX, 4, BLOCK, PRIVATE, EXACT, BOUNDED>;
    if constexpr (cap >= 2)
        if (vec >= 2)
            return fused_tile_kernel<DMAX, 2, BLOCK, PRIVATE, EXACT, BOUNDED>;
    return fused_tile_kernel<DMAX, 1, BLOCK, PRIVATE, EXACT, BOUNDED>;
}

template <int DMAX, bool PRIVATE, bool EXACT, bool BOUNDED>
FusedKernel pick_vec(int vec, bool block256)
{
    if (PRIVATE && block256)
        return pick_vec_block<DMAX, 256, PRIVATE, EXACT, BOUNDED>(vec);
    return pick_vec_block<DMAX, 0, PRIVATE, EXACT, BOUNDED>(vec);
}

template <int DMAX, bool PRIVATE, bool BOUNDED>
FusedKernel pick_exact(bool exact, int vec, bool block256)
{
    return exact ? pick_vec<DMAX, PRIVATE, true, BOUNDED>(vec, block256)
                 : pick_vec<DMAX, PRIVATE, false, BOUNDED>(vec, block256);
}

// BOUNDED (keep-radius shortcut) exists for the private-tile kernel only.
template <bool PRIVATE, bool BOUNDED>
FusedKernel pick_fused(int dmax, bool exact, int vec, bool block256)
{
    static_assert(PRIVATE || !BOUNDED, "bounded pass: private tile only");
    switch (dmax) {
    case 1: return pick_vec<1, PRIVATE, true, BOUNDED>(vec, block256);
    case 2: return pick_exact<2, PRIVATE, BOUNDED>(exact, vec, block256);
    case 4: return pick_exact<4, PRIVATE, BOUNDED>(exact, vec, block256);
    case 8: return pick_exact<8, PRIVATE, BOUNDED>(exact, vec, block256);
    case 16: return pick_exact<16, PRIVATE, BOUNDED>(exact, vec, block256);
    case 32: return pick_exact<32, PRIVATE, BOUNDED>(exact, vec, block256);
    case 64: return pick_exact<64, PRIVATE, BOUNDED>(exact, vec, block256);
    default: return nullptr;
    }
}

}  // namespace

// Common argument checks of the point-streaming ops.
static int check_sizes(uint64_t d, uint64_t n, uint64_t k)
{
    CLK_REQUIRE(d >= 1 && k >= 1, CLK_ERR_SIZE, "d and k must be >= 1");
    CLK_REQUIRE(n < (1ull << 40), CLK_ERR_SIZE, "n too large");
    CLK_REQUIRE(k <= 0xFFFFFFFFull && d <= 0xFFFFull, CLK_ERR_SIZE, "k or d too large");
    return CLK_OK;
}

static int block_size(clk_context *ctx, const clk_kernel_config *cfg)
{
    uint64_t b = cfg && cfg->local_size[0] ? cfg->local_size[0] : 256;
    if (b > (uint64_t)ctx->max_threads_per_block)
        return -1;
    return (int)b;
}

// mode 1: sums/masses accumulate (clk_fused).  mode 2: sums/masses are
// overwritten and new_centroids = divide (clk_lloyd_step).
int run_fused(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n, uint64_t k,
              const float *points, const float *old_centroids, double *sums, uint32_t *labels,
              uint32_t *masses, hipStream_t stream, uint32_t mode, float *new_centroids,
              double *max_shift)
{
    CLK_TRY(check_sizes(d, n, k));
    CLK_REQUIRE(points && old_centroids && sums && labels && masses, CLK_ERR_INVALID_ARGUMENT,
                "fused: null buffer");
    CLK_REQUIRE(mode == 1 || new_centroids, CLK_ERR_INVALID_ARGUMENT, "lloyd_step: null new_centroids");
    CLK_REQUIRE(mode == 1 || new_centroids != old_centroids, CLK_ERR_INVALID_ARGUMENT,
                "lloyd_step: new_centroids must not alias old_centroids");
    const int strategy = cfg ? cfg->strategy : CLK_FUSED_CLUSTER_MERGE;
    CLK_REQUIRE(strategy == CLK_FUSED_CLUSTER_MERGE || strategy == CLK_FUSED_FEATURE_SUM ||
                    strategy == CLK_FUSED_CLUSTER_MERGE_BOUNDED,
                CLK_ERR_INVALID_ARGUMENT, "fused: unknown strategy id");
    const bool bounded = strategy == CLK_FUSED_CLUSTER_MERGE_BOUNDED;
    const int B = block_size(ctx, cfg);
    CLK_REQUIRE(B > 0, CLK_ERR_INVALID_ARGUMENT, "fused: local_size exceeds the device limit");
    const int dmax = dmax_bucket(d);
    const int vec = std::min(pick_vector_width(cfg ? cfg->vector_length : 4, n, points, labels),
                             fused_vec_cap(dmax));
    CLK_REQUIRE(dmax < 64 || B <= 512, CLK_ERR_INVALID_ARGUMENT,
                "fused: local_size above 512 with more than 32 features");

    // Variant by LDS fit (the reference picks its GLOBAL_MEM variant the same
    // way, fused_feature_sum.hpp:315-337): private tile, shared tile, or the
    // separate labeling, mass and centroid-update stages for k*d beyond LDS.
    const uint64_t cen_bytes = dmax ? ((k * dmax + 3) & ~3ull) * 4 : 0;
    const uint64_t rows = k * (d + 1);
    const uint64_t priv_plain = cen_bytes + rows * (uint64_t)B * 4;
    const uint64_t shared_bytes = cen_bytes + rows * 4;
    bool use_private = dmax && strategy != CLK_FUSED_FEATURE_SUM && priv_plain <= ctx->lds_per_block;
    // The bounded pass needs k keep thresholds beside the private tile; where
    // they do not fit it runs as "cluster_merge" (the same results).
    const uint64_t thr_bytes = ((k + 3) & ~3ull) * 4;
    const bool run_bounded = bounded && use_private && priv_plain + thr_bytes <= ctx->lds_per_block;
    const uint64_t priv_bytes = priv_plain + (run_bounded ? thr_bytes : 0);
    bool use_shared = dmax && !use_private && shared_bytes <= ctx->lds_per_block;

    if ((!use_private && !use_shared) || n == 0) {
        if (mode == 2) {
            CLK_TRY(clk_fill_zero(ctx, sums, d * k * sizeof(double), stream));
            CLK_TRY(clk_fill_zero(ctx, masses, k * sizeof(uint32_t), stream));
        }
        if (n) {
            // labels, then the fixed-order mass histogram and centroid sums
            // (feature slices for large k*d; global atomics only past those)
            CLK_TRY(launch_labeling(ctx, cfg, d, n, k, points, old_centroids, labels, stream));
            CLK_TRY(run_mass_update(ctx, nullptr, n, k, labels, masses, stream));
            CLK_TRY(run_centroid_update(ctx, nullptr, d, n, k, points, sums, labels, stream));
        }
        if (mode == 2)
            return clk_matrix_divide_row(ctx, d, k, sums, masses, new_centroids,
                                         max_shift ? old_centroids : nullptr, max_shift, stream);
        return CLK_OK;
    }

    const bool b256 = (B == 256);
    const bool exact = (uint64_t)dmax == d;
    FusedKernel kern = !use_private ? pick_fused<false, false>(dmax, exact, vec, b256)
                       : run_bounded ? pick_fused<true, true>(dmax, exact, vec, b256)
                                     : pick_fused<true, false>(dmax, exact, vec, b256);
    const size_t lds = use_private ? priv_bytes : shared_bytes;
    const uint64_t groups = n / vec;
    // The bounded pass takes the grid of the exhaustive private-tile kernel:
    // the same points per work-item in the same order, so bit-identical sums.
    const void *grid_kern = run_bounded ? (const void *)pick_fused<true, false>(dmax, exact, vec, b256)
                                        : (const void *)kern;
    const size_t grid_lds = run_bounded ? lds - thr_bytes : lds;
    uint64_t grid = (cfg && cfg->global_size[0]) ? div_up(cfg->global_size[0], (uint64_t)B)
                                                 : auto_grid(ctx, grid_kern, B, grid_lds, groups);
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, 1u << 20));

    CLK_TRY(ensure_scratch(ctx->part_sums, grid * d * k * sizeof(double)));
    CLK_TRY(ensure_scratch(ctx->part_masses, grid * k * sizeof(uint32_t)));
    FinalizeArgs fa;
    fa.part_sums = static_cast<double *>(ctx->part_sums.ptr);
    fa.part_masses = static_cast<uint32_t *>(ctx->part_masses.ptr);
    fa.rows = (uint32_t)grid;
    fa.sums = sums;
    fa.masses = masses;
    fa.centroids = new_centroids;
    fa.old_centroids = max_shift ? old_centroids : nullptr;
    fa.max_shift = max_shift;
    fa.d = (uint32_t)d;
    fa.k = (uint32_t)k;
    fa.mode = mode;
    {
        KernelTimer t(ctx,
                      !use_private ? "FusedFeatureSum"
                      : run_bounded ? "FusedClusterMergeBounded"
                                   : "FusedClusterMerge",
                      stream);
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(B), lds, stream, points, n, (uint32_t)d,
                           (uint32_t)k, old_centroids, labels, static_cast<double *>(ctx->part_sums.ptr),
                           static_cast<uint32_t *>(ctx->part_masses.ptr));
        CLK_HIP_CHECK(hipGetLastError());
    }
    const size_t fin_lds = d * k * sizeof(double) + k * sizeof(uint32_t) + 12 * 1024 + 16;
    if (fin_lds <= ctx->lds_per_block)
        return launch_finalize(ctx, fa, stream);
    // d*k beyond the one-work-group merge: column reduction, then divide.
    if (mode == 2) {
        CLK_TRY(clk_fill_zero(ctx, sums, d * k * sizeof(double), stream));
        CLK_TRY(clk_fill_zero(ctx, masses, k * sizeof(uint32_t), stream));
    }
    CLK_TRY(launch_reduce_partials(ctx, fa.part_sums, d * k, sums, fa.part_masses, k, masses, grid,
                                   stream));
    if (mode == 2)
        return clk_matrix_divide_row(ctx, d, k, sums, masses, new_centroids,
                                     max_shift ? old_centroids : nullptr, max_shift, stream);
    return CLK_OK;
}

}  // namespace clk

extern "C" int clk_fused(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                         uint64_t k, const float *points, const float *old_centroids,
                         double *sums, uint32_t *labels, uint32_t *masses, void *stream)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(!cfg || cfg->stage == CLK_STAGE_FUSED, CLK_ERR_INVALID_ARGUMENT,
                "clk_fused: config is not a fused-stage config");
    return clk::run_fused(ctx, cfg, d, n, k, points, old_centroids, sums, labels, masses,
                          clk::as_stream(stream), 1, nullptr, nullptr);
}

extern "C" int clk_lloyd_step(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                              uint64_t n, uint64_t k, const float *points,
                              const float *old_centroids, float *new_centroids, double *sums,
                              uint32_t *labels, uint32_t *masses, double *max_shift, void *stream)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(!cfg || cfg->stage == CLK_STAGE_FUSED, CLK_ERR_INVALID_ARGUMENT,
                "clk_lloyd_step: config is not a fused-stage config");
    return clk::run_fused(ctx, cfg, d, n, k, points, old_centroids, sums, labels, masses,
                          clk::as_stream(stream), 2, new_centroids, max_shift);
}
