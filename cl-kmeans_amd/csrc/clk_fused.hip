// Fused single pass (labeling + mass histogram + centroid feature sums in one
// read of the points): the work of lloyd_fused_cluster_merge.cl and
// lloyd_fused_feature_sum.cl, designed for gfx950.
//
// Layout per work-group (LDS):
//   cen   [k][DMAX]            centroids, zero-padded to the register bucket
//   thr   [k]                  keep thresholds (bounded pass)
//   scl   [DMAX]               grid scales of the features (clk_grid.hpp)
//   acc   [d][k] int64         exact grid-integer sums of the work-group
//   flg   [d][k]               off-grid flags
//   tile  cluster_merge: [k*(d+1)][B] uint32 -- row c*(d+1) = mass of c,
//                        rows c*(d+1)+1+f = int32 grid sums of feature f,
//                        one column per work-item (no collisions);
//         feature_sum:   [k] uint32 masses (the sums go to acc directly).
// Each work-item labels VEC consecutive points per step with 128-bit loads
// (one per feature row), writes the labels, rounds the coordinates to their
// grid integers and adds them: cluster_merge into its own int32 column
// (ds_add_u32), flushed into acc (ds_add_u64) every 64 values a cell can
// take; feature_sum straight into acc.  All additions are exact integer
// additions, so the sums do not depend on the grid or the order.  At the end
// acc (times the unit) and the summed mass columns are written as
// per-work-group FP64 / uint32 partials; a column reduction adds the
// partials into the caller's sums and masses (exact: every partial is a
// multiple of the unit below 2^53 units).
#include "clk_common.hpp"
#include "clk_finalize.hpp"
#include "clk_grid.hpp"
#include "clk_kernels.hpp"

#include <algorithm>
#include <mutex>

namespace clk {
namespace {

__host__ __device__ constexpr uint32_t align4(uint32_t v) { return (v + 3u) & ~3u; }

// LDS floats before the tile (host and device agree on the layout).
__host__ __device__ inline uint32_t fused_head_words(uint32_t d, uint32_t k, uint32_t dmax, bool bounded)
{
    return align4(k * dmax) + (bounded ? align4(k) : 0u) + align4(dmax) + 2u * d * k + align4(d * k);
}

template <int DMAX, int VEC, int BLOCK, bool PRIVATE, bool EXACT, bool BOUNDED>
__global__ void __launch_bounds__(BLOCK > 0 ? BLOCK : (DMAX >= 64 ? 512 : 1024))
fused_tile_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                  const float *__restrict__ cen, uint32_t *__restrict__ labels, GridRef grid,
                  double *__restrict__ part_sums, uint32_t *__restrict__ part_masses)
{
    extern __shared__ __align__(16) float smem[];
    const uint32_t B = BLOCK > 0 ? BLOCK : blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t cols = PRIVATE ? B : 1;
    const uint32_t dk = d * k;
    const uint32_t rows = PRIVATE ? k * (d + 1) : k;
    float *cen_lds = smem;
    float *thr_lds = smem + align4(k * DMAX);  // BOUNDED: keep thresholds [k]
    float *scl = thr_lds + (BOUNDED ? align4(k) : 0u);
    long long *acc = reinterpret_cast<long long *>(scl + align4(DMAX));  // 16-byte aligned
    uint32_t *flg = reinterpret_cast<uint32_t *>(acc + dk);
    uint32_t *tile = flg + align4(dk);

    stage_centroids<DMAX>(cen, d, k, cen_lds);
    for (uint32_t f = tid; f < DMAX; f += B)
        scl[f] = f < d ? grid_scale(grid.gmax[f], grid.bits) : 1.0f;
    for (uint32_t i = tid; i < dk; i += B) {
        acc[i] = 0;
        flg[i] = 0u;
    }
    for (uint32_t i = tid; i < rows * cols; i += B)
        tile[i] = 0u;
    __syncthreads();
    if constexpr (BOUNDED) {
        stage_keep_thresholds<DMAX>(cen_lds, k, thr_lds);
        __syncthreads();
    }
    // grid scales in registers for small buckets, read from LDS otherwise
    constexpr bool kRegScale = DMAX <= 16;
    float sc_reg[kRegScale ? DMAX : 1];
    if constexpr (kRegScale) {
#pragma unroll
        for (int f = 0; f < DMAX; ++f)
            sc_reg[f] = scl[f];
    }
    auto scale_of = [&](int f) -> float {
        if constexpr (kRegScale)
            return sc_reg[f];
        else
            return scl[f];
    };

    // the points' own grid with no NaN / inf in them: every coordinate is on
    // the grid (no range checks, no flags)
    const bool on_grid = grid.own != 0 && grid.gmax[d] == 0;
    const uint64_t groups = n / VEC;
    const uint64_t stride = (uint64_t)gridDim.x * B;
    uint32_t *mine = tile + (PRIVATE ? tid : 0);
    const uint32_t row_stride = (d + 1) * cols;  // < 2^24: the tile fits in LDS
    // A private cell takes at most VEC values per step and holds 64 values of
    // |q| < 2^25 without int32 overflow: flush every 64 / VEC steps.
    constexpr uint32_t kFlushSteps = 64 / VEC;
    uint32_t since = 0;
    auto flush = [&]() {
        // clusters in a rotated order so the lanes of a wave hit different
        // acc cells
        uint32_t c = tid % k;
        for (uint32_t i = 0; i < k; ++i) {
            uint32_t *cell = mine + __umul24(c, row_stride) + cols;
#pragma unroll 1
            for (uint32_t f = 0; f < d; ++f, cell += cols) {
                const int32_t v = (int32_t)*cell;
                *cell = 0u;
                lds_add64(acc + f * k + c, (long long)v);
            }
            c = (c + 1 == k) ? 0u : c + 1;
        }
    };
    // Software pipeline, two register buffers: one step's points are in
    // flight while the other step's are labelled and accumulated.  A step
    // takes V = VEC points p .. p+V-1 (V = 1 for the n % VEC tail points).
    auto step = [&](const auto &x, const auto &prev, uint64_t p) {
        constexpr int V = (int)(sizeof(prev) / sizeof(uint32_t));
        auto put = [&](const uint32_t (&lab)[V]) { store_labels<V>(labels, p, lab); };
        uint32_t lab[V];
        if constexpr (BOUNDED) {
            // Keep the previous label where the point is inside its centroid's
            // keep radius (keep_threshold: the full scan provably returns it);
            // otherwise scan every centroid.  Only changed labels are stored.
            bool keep = true;
#pragma unroll
            for (int v = 0; v < V; ++v) {
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
                argmin_lds<DMAX, V>(x, cen_lds, k, lab);
            bool changed = false;
#pragma unroll
            for (int v = 0; v < V; ++v)
                changed = changed || lab[v] != prev[v];
            if (changed)
                put(lab);
        } else {
            argmin_lds<DMAX, V>(x, cen_lds, k, lab);
            put(lab);
        }
        bool off = false;
        if (on_grid) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                uint32_t *t = mine + __umul24(lab[v], row_stride);
                lds_add(PRIVATE ? t : tile + lab[v], 1u);
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (EXACT || f < (int)d) {
                        const int32_t q = grid_q_on(x[f][v], scale_of(f));
                        if (PRIVATE)
                            lds_add(t + (f + 1) * cols, (uint32_t)q);
                        else
                            lds_add64(acc + (uint32_t)f * k + lab[v], (long long)q);
                    }
            }
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                uint32_t *t = mine + __umul24(lab[v], row_stride);
                // u32 LDS atomics run at full speed (ds_add_f32 does not)
                lds_add(PRIVATE ? t : tile + lab[v], 1u);
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (EXACT || f < (int)d) {
                        bool ok;
                        const int32_t q = grid_q(x[f][v], scale_of(f), ok);
                        off = off || !ok;
                        if (PRIVATE)
                            lds_add(t + (f + 1) * cols, (uint32_t)q);
                        else
                            lds_add64(acc + (uint32_t)f * k + lab[v], (long long)q);
                    }
            }
        }
        if (__builtin_expect(__ballot(off) != 0, 0)) {
            // values off the grid (NaN, inf, outside a pinned grid) were
            // added as 0; their cells are flagged
            if (off) {
#pragma unroll
                for (int v = 0; v < V; ++v)
#pragma unroll
                    for (int f = 0; f < DMAX; ++f)
                        if ((EXACT || f < (int)d) &&
                            !(__builtin_fabsf(x[f][v] * scale_of(f)) < kGridLimit))
                            lds_or(flg + (uint32_t)f * k + lab[v], special_flag(x[f][v]));
            }
        }
        if (PRIVATE && ++since == kFlushSteps) {
            flush();
            since = 0;
        }
    };
    // Loads are unconditional (index clamped to the last group, always in
    // bounds) so the loop body is straight-line and the wait counters can
    // keep the prefetch in flight across the labeling of the other buffer.
    // BOUNDED also prefetches the previous labels of the same points.
    // (Buffer-resource loads with 32-bit offsets ran 4% slower than these
    // global loads at N=2^27, d=4, k=10.)
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
        step(xa, la, g * VEC);
        if (g1 >= groups)
            break;
        const uint64_t g2 = g1 + stride;
        load(min(g2, last), xa, la);
        step(xb, lb, g1 * VEC);
        g = g2;
    }
    if constexpr (VEC > 1) {
        // the n % VEC points after the last group, one at a time, by the
        // work-item that would take group `groups` next
        if ((uint64_t)blockIdx.x * B + tid == groups % stride) {
            for (uint64_t p = groups * VEC; p < n; ++p) {
                float x1[DMAX][1];
                uint32_t l1[1] = {0u};
                load_points<DMAX, 1, EXACT>(pts, n, d, p, x1);
                if constexpr (BOUNDED)
                    stream_load_labels<1>(labels, p, l1);
                step(x1, l1, p);
            }
        }
    }
    if (PRIVATE && since != 0)
        flush();
    __syncthreads();

    for (uint32_t r = tid; r < k * (d + 1); r += B) {
        // r < k: masses; then the d*k sums in [d][k] order
        if (r < k) {
            part_masses[(uint64_t)blockIdx.x * k + r] =
                PRIVATE ? lds_row_sum_u32(tile + (uint64_t)r * (d + 1) * cols, cols, r) : tile[r];
        } else {
            const uint32_t i = r - k, f = i / k;
            part_sums[(uint64_t)blockIdx.x * dk + i] = cell_value(acc[i], grid_unit(scl[f]), flg[i]);
        }
    }
}

using FusedKernel = void (*)(const float *, uint64_t, uint32_t, uint32_t, const float *,
                             uint32_t *, GridRef, double *, uint32_t *);

// Points per work-item at most (VEC x DMAX values in registers, no spills).
constexpr int fused_vec_cap(int dmax) { return dmax >= 32 ? 1 : dmax >= 16 ? 2 : 4; }

template <int DMAX, int BLOCK, bool PRIVATE, bool EXACT, bool BOUNDED>
FusedKernel pick_vec_block(int vec)
{
    // work-groups of any size (up to 1024 work-items, 128 VGPRs): half the
    // points per work-item from 8 features on (no scratch)
    constexpr int cap = BLOCK > 0 ? fused_vec_cap(DMAX) : fused_vec_cap(2 * DMAX);
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
// Caller-provided device scratch (clk_lloyd_step_noalloc): partial rows and
// the per-call sum grid are carved from it instead of the context's
// allocations.
struct ExtScratch {
    char *ptr;
    size_t bytes;
};

// grid_in: the sum grid to use (the batch engine's per-job grid) instead of
// the pinned or per-call one; only the LDS-tile paths take it.
int run_fused(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n, uint64_t k,
              const float *points, const float *old_centroids, double *sums, uint32_t *labels,
              uint32_t *masses, hipStream_t stream, uint32_t mode, float *new_centroids,
              double *max_shift, const ExtScratch *ext = nullptr, const GridRef *grid_in = nullptr)
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
    const int vec = std::min(pick_vector_width(cfg ? cfg->vector_length : 4), fused_vec_cap(dmax));
    CLK_REQUIRE(dmax < 64 || B <= 512, CLK_ERR_INVALID_ARGUMENT,
                "fused: local_size above 512 with more than 32 features");

    // Variant by LDS fit (the reference picks its GLOBAL_MEM variant the same
    // way, fused_feature_sum.hpp:315-337): private tile, shared tile, or the
    // separate labeling, mass and centroid-update stages for k*d beyond LDS.
    const uint64_t rows = k * (d + 1);
    const uint64_t head_plain = dmax ? (uint64_t)fused_head_words((uint32_t)d, (uint32_t)k, dmax, false) * 4 : 0;
    const uint64_t thr_bytes = ((k + 3) & ~3ull) * 4;
    const uint64_t priv_plain = head_plain + rows * (uint64_t)B * 4;
    const uint64_t shared_bytes = head_plain + thr_bytes;  // masses [k]
    bool use_private = dmax && strategy != CLK_FUSED_FEATURE_SUM && k * d < (1ull << 24) &&
                       priv_plain <= ctx->lds_per_block;
    // The bounded pass needs k keep thresholds beside the private tile; where
    // they do not fit it runs as "cluster_merge" (the same results).
    const bool run_bounded = bounded && use_private && priv_plain + thr_bytes <= ctx->lds_per_block;
    const uint64_t priv_bytes = priv_plain + (run_bounded ? thr_bytes : 0);
    bool use_shared = dmax && !use_private && k * d < (1ull << 24) && shared_bytes <= ctx->lds_per_block;

    CLK_REQUIRE(!ext || use_private || use_shared || n == 0, CLK_ERR_UNSUPPORTED,
                "lloyd_step_noalloc: k*d beyond the work-group LDS tiles (the separate stages allocate)");
    CLK_REQUIRE(!grid_in || use_private || use_shared || n == 0, CLK_ERR_UNSUPPORTED,
                "fused pass with a given grid: k*d beyond the work-group LDS tiles");
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
    uint64_t grid = (cfg && cfg->global_size[0]) ? div_up(cfg->global_size[0], (uint64_t)B)
                                                 : auto_grid(ctx, (const void *)kern, B, lds, groups);
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, 1u << 20));
    if (lds > 65536)
        CLK_HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    GridRef sgrid;
    double *psums = nullptr;
    uint32_t *pmasses = nullptr;
    if (ext) {
        // [per-call grid d words, 256-aligned] [partial sums] [partial masses]:
        // as many work-groups as the rows that fit (grid-stride: any count
        // gives the same results)
        const size_t head = ctx->grid_pinned ? 0 : ((d + 1) * sizeof(uint32_t) + 255) / 256 * 256;
        const size_t row = d * k * sizeof(double) + k * sizeof(uint32_t);
        const size_t fit = ext->bytes > head ? (ext->bytes - head) / row : 0;
        CLK_REQUIRE(fit >= 1, CLK_ERR_SIZE, "lloyd_step_noalloc: k*(8d+4) exceeds the static scratch");
        grid = std::min<uint64_t>(grid, fit);
        if (ctx->grid_pinned) {
            CLK_TRY(resolve_grid(ctx, d, n, points, stream, &sgrid));
        } else {
            uint32_t *gmax = reinterpret_cast<uint32_t *>(ext->ptr);
            CLK_TRY(launch_grid_max(ctx, d, n, points, gmax, stream));
            sgrid.gmax = gmax;
            sgrid.bits = grid_bits(n);
            sgrid.own = 1;
        }
        psums = reinterpret_cast<double *>(ext->ptr + head);
        pmasses = reinterpret_cast<uint32_t *>(psums + grid * d * k);
    } else {
        if (grid_in)
            sgrid = *grid_in;
        else
            CLK_TRY(resolve_grid(ctx, d, n, points, stream, &sgrid));
        CLK_TRY(ensure_scratch(ctx->part_sums, grid * d * k * sizeof(double)));
        CLK_TRY(ensure_scratch(ctx->part_masses, grid * k * sizeof(uint32_t)));
        psums = static_cast<double *>(ctx->part_sums.ptr);
        pmasses = static_cast<uint32_t *>(ctx->part_masses.ptr);
    }
    FinalizeArgs fa;
    fa.part_sums = psums;
    fa.part_masses = pmasses;
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
                           (uint32_t)k, old_centroids, labels, sgrid, psums, pmasses);
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

int lloyd_step_with_grid(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n, uint64_t k,
                         const float *points, const float *old_centroids, float *new_centroids, double *sums,
                         uint32_t *labels, uint32_t *masses, double *max_shift, const GridRef &grid,
                         hipStream_t stream)
{
    return run_fused(ctx, cfg, d, n, k, points, old_centroids, sums, labels, masses, stream, 2, new_centroids,
                     max_shift, nullptr, &grid);
}

// ---- the entry point that allocates nothing -----------------------------

namespace {

// Scratch of clk_lloyd_step_noalloc: partial rows and the per-call sum grid.
// Every GPU the library runs on holds one copy until the process ends.  In
// host code the name is a host-side stand-in: kernels and fills take the
// device address hipGetSymbolAddress returns (scratch_dev below).
static __device__ __align__(256) unsigned char scratch[CLK_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_dev[kMaxDevices] = {};

}  // namespace

// Once per device, with that device current (clk_context_create calls it):
// the device address of this device's copy of `scratch`.
int noalloc_setup(clk_context *ctx)
{
    CLK_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices, CLK_ERR_UNSUPPORTED,
                "noalloc scratch: device index beyond the table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[ctx->device]) {
        void *p = nullptr;
        CLK_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CLK_REQUIRE(p != nullptr, CLK_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_dev[ctx->device] = static_cast<char *>(p);
    }
    ctx->noalloc_scratch = g_scratch_dev[ctx->device];
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

extern "C" int clk_lloyd_step_noalloc(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                                      uint64_t n, uint64_t k, const float *points,
                                      const float *old_centroids, float *new_centroids, double *sums,
                                      uint32_t *labels, uint32_t *masses, double *max_shift,
                                      void *stream)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(!cfg || cfg->stage == CLK_STAGE_FUSED, CLK_ERR_INVALID_ARGUMENT,
                "clk_lloyd_step_noalloc: config is not a fused-stage config");
    CLK_REQUIRE(ctx->noalloc_scratch != nullptr, CLK_ERR_HIP, "clk_lloyd_step_noalloc: no scratch address");
    const clk::ExtScratch ext{ctx->noalloc_scratch, CLK_NOALLOC_SCRATCH_BYTES};
    // no timing records: taking an event may create one
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    const int rc = clk::run_fused(ctx, cfg, d, n, k, points, old_centroids, sums, labels, masses,
                                  clk::as_stream(stream), 2, new_centroids, max_shift, &ext);
    ctx->profiling = prof;
    return rc;
}
