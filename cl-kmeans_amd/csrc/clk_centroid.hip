// Centroid update: sums[f,c] += sum_{p : labels[p] == c} x_f, the three
// strategies of CentroidUpdateFactory built for gfx950:
//   cluster_merge       one LDS tile column [k*d] per work-item (no
//                       collisions), all features of the work-item's points
//                       (lloyd_cluster_merge.cl's private k x d tile)
//   feature_sum_pardim  2-D grid: blockIdx.y selects a slice of
//                       local_features*thread_features features, threadIdx.y
//                       a thread_features sub-slice; private LDS columns
//                       (lloyd_feature_sum_pardim.cl's points x feature tiles)
//   feature_sum         one LDS accumulator [d][k] shared by the work-group,
//                       work-items stride over the flattened feature-major
//                       matrix (coalesced rows); the smallest LDS footprint
// Every coordinate is rounded to its grid integer and added exactly
// (clk_grid.hpp): the private int32 columns flush into the work-group's
// int64 accumulator (ds_add_u64) at least every 64 values, shared
// accumulators take ds_add_u64 directly.  The work-group's sums (times the
// unit) form its FP64 partial row; a column reduction adds the rows (exact).
// So every strategy and geometry gives the same bits.  For large k*d (when
// neither fits) every strategy uses the cluster-major slice kernel below;
// the FP64 global-atomic pass (of grid values, so exact as well) remains for
// k too large for that.  `masses` is not read, as in the reference kernels.
#include "clk_common.hpp"
#include "clk_grid.hpp"
#include "clk_kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace clk {
namespace {

__host__ __device__ constexpr uint32_t align4u(uint32_t v) { return (v + 3u) & ~3u; }

// LDS words before a kernel's tile: scales [nf] | acc [nf][k] int64 |
// flags [nf][k].
__host__ __device__ inline uint32_t acc_head_words(uint32_t nf, uint32_t k)
{
    return align4u(nf) + 2u * nf * k + align4u(nf * k);
}

struct AccHead {
    float *scl;
    long long *acc;
    uint32_t *flg;
    uint32_t *tail;
};

// Carve the head; stage the scales of features f0 .. f0+nf-1 and zero the
// accumulator and flags (callers add a barrier).
__device__ inline AccHead acc_head(void *lds, uint32_t nf, uint32_t k, uint32_t f0, uint32_t d,
                                   const GridRef &grid, uint32_t tid, uint32_t nthreads)
{
    AccHead h;
    h.scl = static_cast<float *>(lds);
    h.acc = reinterpret_cast<long long *>(h.scl + align4u(nf));
    h.flg = reinterpret_cast<uint32_t *>(h.acc + (size_t)nf * k);
    h.tail = h.flg + align4u(nf * k);
    for (uint32_t j = tid; j < nf; j += nthreads)
        h.scl[j] = f0 + j < d ? grid_scale(grid.gmax[f0 + j], grid.bits) : 1.0f;
    for (uint32_t i = tid; i < nf * k; i += nthreads) {
        h.acc[i] = 0;
        h.flg[i] = 0u;
    }
    return h;
}

// Add x (feature slot j of the head) of a point of cluster c into an int32
// private cell, flagging off-grid values.
__device__ __forceinline__ void add_private(uint32_t *cell, float x, float scale, const AccHead &h,
                                            uint32_t j, uint32_t k, uint32_t c)
{
    bool ok;
    const int32_t q = grid_q(x, scale, ok);
    lds_add(cell, (uint32_t)q);
    if (__builtin_expect(!ok, 0))
        lds_or(h.flg + j * k + c, special_flag(x));
}

template <int DMAX, int VEC, int BLOCK>
__global__ void __launch_bounds__(BLOCK > 0 ? BLOCK : 1024)
cu_cluster_merge_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                        const uint32_t *__restrict__ labels, GridRef grid, double *__restrict__ part)
{
    extern __shared__ __align__(16) float lds_cm[];
    const uint32_t B = BLOCK > 0 ? BLOCK : blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t rows = k * d;
    const AccHead h = acc_head(lds_cm, d, k, 0, d, grid, tid, B);
    uint32_t *tile = h.tail;  // [k*d][B] int32 grid sums
    for (uint32_t i = tid; i < rows * B; i += B)
        tile[i] = 0u;
    __syncthreads();
    uint32_t *mine = tile + tid;
    const uint32_t row_stride = d * B;
    const uint64_t groups = n / VEC;
    const uint64_t stride = (uint64_t)gridDim.x * B;
    // a cell takes at most VEC values per step: flush every 64 / VEC steps
    constexpr uint32_t kFlushSteps = 64 / VEC;
    uint32_t since = 0;
    auto flush = [&]() {
        uint32_t c = tid % k;
        for (uint32_t i = 0; i < k; ++i) {
            uint32_t *cell = mine + __umul24(c, row_stride);
#pragma unroll 1
            for (uint32_t f = 0; f < d; ++f, cell += B) {
                const int32_t v = (int32_t)*cell;
                *cell = 0u;
                lds_add64(h.acc + f * k + c, (long long)v);
            }
            c = (c + 1 == k) ? 0u : c + 1;
        }
    };
    // grid scales in registers for small buckets, read from LDS otherwise
    constexpr bool kRegScale = DMAX <= 16;
    float sc_reg[kRegScale ? DMAX : 1];
    if constexpr (kRegScale) {
#pragma unroll
        for (int f = 0; f < DMAX; ++f)
            sc_reg[f] = h.scl[f];
    }
    auto scale_of = [&](int f) -> float {
        if constexpr (kRegScale)
            return sc_reg[f];
        else
            return h.scl[f];
    };
    // the points' own grid with no NaN / inf in them: every coordinate is on
    // the grid (no range checks, no flags)
    const bool on_grid = grid.own != 0 && grid.gmax[d] == 0;
    // V points p .. p+V-1 (V = VEC, or 1 for the n % VEC tail points)
    auto add_points = [&](const auto &x, const auto &lab) {
        constexpr int V = (int)(sizeof(lab) / sizeof(uint32_t));
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (lab[v] >= k)
                continue;
            uint32_t *t = mine + __umul24(lab[v], row_stride);
            if (on_grid) {
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (f < (int)d)
                        lds_add(t + f * B, (uint32_t)grid_q_on(x[f][v], scale_of(f)));
            } else {
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (f < (int)d)
                        add_private(t + f * B, x[f][v], scale_of(f), h, (uint32_t)f, k, lab[v]);
            }
        }
        if (++since == kFlushSteps) {
            flush();
            since = 0;
        }
    };
    auto load = [&](uint64_t gg, float (&x)[DMAX][VEC], uint32_t (&l)[VEC]) {
        load_points<DMAX, VEC>(pts, n, d, gg * VEC, x);
        load_labels<VEC>(labels, gg * VEC, l);
    };
    uint64_t g = (uint64_t)blockIdx.x * B + tid;
    if constexpr (DMAX * VEC <= 32) {
        // Software pipeline, two register buffers: one step's points and
        // labels are in flight while the other step's are accumulated.
        // Loads are unconditional (index clamped to the last group, always
        // in bounds).  (Wider steps would not fit the registers twice.)
        const uint64_t last = groups ? groups - 1 : 0;
        float xa[DMAX][VEC], xb[DMAX][VEC];
        uint32_t la[VEC], lb[VEC];
        if (groups)
            load(min(g, last), xa, la);
        while (g < groups) {
            const uint64_t g1 = g + stride;
            load(min(g1, last), xb, lb);
            add_points(xa, la);
            if (g1 >= groups)
                break;
            const uint64_t g2 = g1 + stride;
            load(min(g2, last), xa, la);
            add_points(xb, lb);
            g = g2;
        }
    } else {
        for (; g < groups; g += stride) {
            float x[DMAX][VEC];
            uint32_t l[VEC];
            load(g, x, l);
            add_points(x, l);
        }
    }
    if constexpr (VEC > 1) {
        // the n % VEC points after the last group, one at a time, by the
        // work-item that would take group `groups` next
        if ((uint64_t)blockIdx.x * B + tid == groups % stride) {
            for (uint64_t p = groups * VEC; p < n; ++p) {
                float x1[DMAX][1];
                uint32_t l1[1];
                load_points<DMAX, 1>(pts, n, d, p, x1);
                load_labels<1>(labels, p, l1);
                add_points(x1, l1);
            }
        }
    }
    if (since != 0)
        flush();
    __syncthreads();
    for (uint32_t i = tid; i < rows; i += B) {
        const uint32_t f = i / k;  // i = f*k + c
        part[(uint64_t)blockIdx.x * rows + i] = cell_value(h.acc[i], grid_unit(h.scl[f]), h.flg[i]);
    }
}

// blockDim = (Bx, lf); block (x, y) owns features [y*FS, y*FS+FS), FS = lf*tf;
// thread (tx, ty) owns features y*FS + ty*tf + [0, tf) of points
// x*Bx + tx + i*gridDim.x*Bx.  Private LDS rows c*FS + (ty*tf + j), columns tx.
__global__ void __launch_bounds__(1024)
cu_pardim_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k, uint32_t tf,
                 const uint32_t *__restrict__ labels, GridRef grid, double *__restrict__ part)
{
    extern __shared__ __align__(16) float lds_pd[];
    const uint32_t Bx = blockDim.x, lf = blockDim.y;
    const uint32_t FS = lf * tf;
    const uint32_t f0 = blockIdx.y * FS;
    const uint32_t tlin = threadIdx.y * Bx + threadIdx.x;
    const uint32_t nthreads = Bx * lf;
    const uint32_t rows = k * FS;
    const AccHead h = acc_head(lds_pd, FS, k, f0, d, grid, tlin, nthreads);
    uint32_t *tile = h.tail;
    for (uint32_t i = tlin; i < rows * Bx; i += nthreads)
        tile[i] = 0u;
    __syncthreads();
    const uint32_t fsub = threadIdx.y * tf;  // first feature of this thread in the slice
    const uint64_t stride = (uint64_t)gridDim.x * Bx;
    // a cell takes at most one value per point: flush every 64 points
    uint32_t since = 0;
    auto flush = [&]() {
        for (uint32_t c = 0; c < k; ++c) {
            uint32_t *t = tile + ((uint64_t)c * FS + fsub) * Bx + threadIdx.x;
            for (uint32_t j = 0; j < tf; ++j) {
                const int32_t v = (int32_t)t[j * Bx];
                t[j * Bx] = 0u;
                lds_add64(h.acc + (fsub + j) * k + c, (long long)v);
            }
        }
    };
    for (uint64_t p = (uint64_t)blockIdx.x * Bx + threadIdx.x; p < n; p += stride) {
        const uint32_t c = labels[p];
        if (c < k) {
            uint32_t *t = tile + ((uint64_t)c * FS + fsub) * Bx + threadIdx.x;
            for (uint32_t j = 0; j < tf; ++j) {
                const uint32_t f = f0 + fsub + j;
                if (f < d)
                    add_private(t + j * Bx, pts[(uint64_t)f * n + p], h.scl[fsub + j], h, fsub + j, k, c);
            }
        }
        if (++since == 64) {
            flush();
            since = 0;
        }
    }
    if (since != 0)
        flush();
    __syncthreads();
    for (uint32_t i = tlin; i < FS * k; i += nthreads) {
        const uint32_t j = i / k, c = i % k, f = f0 + j;
        if (f < d)
            part[(uint64_t)blockIdx.x * ((uint64_t)d * k) + (uint64_t)f * k + c] =
                cell_value(h.acc[i], grid_unit(h.scl[j]), h.flg[i]);
    }
}

// Shared accumulator [d][k]; work-items stride over elements e = f*n + p.
__global__ void __launch_bounds__(1024)
cu_feature_sum_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                      const uint32_t *__restrict__ labels, GridRef grid, double *__restrict__ part)
{
    extern __shared__ __align__(16) float lds_fs[];
    const uint32_t rows = d * k;
    const AccHead h = acc_head(lds_fs, d, k, 0, d, grid, threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t total = n * d;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint64_t f = e / n, p = e - f * n;
        const uint32_t c = labels[p];
        if (c < k) {
            const float x = pts[e];
            bool ok;
            const int32_t q = grid_q(x, h.scl[f], ok);
            lds_add64(h.acc + f * k + c, (long long)q);
            if (__builtin_expect(!ok, 0))
                lds_or(h.flg + f * k + c, special_flag(x));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < rows; i += blockDim.x)
        part[(uint64_t)blockIdx.x * rows + i] = cell_value(h.acc[i], grid_unit(h.scl[i / k]), h.flg[i]);
}

// Cluster-major feature slices for large k*d.  blockIdx.y selects features
// [4*blockIdx.y, +4); the work-group shares one int64 accumulator [k][4] in
// LDS.  A wave takes 64 consecutive points (lane = point: 4 coalesced
// feature loads and the label) and adds each point's grid integers into its
// cluster's row with ds_add_u64 (exact, so no ordering between lanes or
// waves is needed).  The accumulator (times the units) is the work-group's
// partial row.
constexpr uint32_t kSliceF = 4, kSliceWaves = 4, kSlicePrefetch = 8;

size_t slice_lds_bytes(uint64_t k)
{
    return (size_t)acc_head_words(kSliceF, (uint32_t)k) * 4;
}

__global__ void __launch_bounds__(64 * kSliceWaves)
cu_slice_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                const uint32_t *__restrict__ labels, GridRef grid, double *__restrict__ part)
{
    extern __shared__ __align__(16) float lds_sl[];
    typedef float f4 __attribute__((ext_vector_type(4)));
    // the wave index as a scalar: tile indices and the loop below are then
    // wave-uniform, so the prefetch ring needs no divergent control flow
    const uint32_t tid = threadIdx.x, lane = tid % 64;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid / 64));
    const uint32_t f0 = blockIdx.y * kSliceF;
    const uint32_t nf = min(kSliceF, d - f0);
    const AccHead h = acc_head(lds_sl, kSliceF, k, f0, d, grid, tid, blockDim.x);
    __syncthreads();
    float sc[kSliceF];
#pragma unroll
    for (uint32_t j = 0; j < kSliceF; ++j)
        sc[j] = h.scl[j];
    const uint64_t tiles = (n + 63) / 64;
    const uint64_t tstride = (uint64_t)gridDim.x * kSliceWaves;
    // Every load is unconditional and in bounds (a tile past the end reads
    // the last tile, a feature past the slice reads its last feature; such
    // values are never added), so no branch separates a load from its use
    // and the loads of later tiles stay in flight.
    auto load = [&](uint64_t tile, f4 &v, uint32_t &L) {
        const uint64_t p = min(min(tile, tiles - 1) * 64 + lane, n - 1);
#pragma unroll
        for (uint32_t j = 0; j < kSliceF; ++j)
            v[j] = pts[(uint64_t)(f0 + min(j, nf - 1)) * n + p];
        L = labels[p];
    };
    // A ring of kSlicePrefetch tiles in registers: the loads of the next
    // kSlicePrefetch tiles of this wave are in flight while one is added
    // (one tile alone leaves too few bytes in flight to cover HBM latency).
    // The ring is indexed by the unrolled slot, so it stays in registers.
    f4 vr[kSlicePrefetch];
    uint32_t Lr[kSlicePrefetch];
    const uint64_t first = (uint64_t)blockIdx.x * kSliceWaves + wave;
    // this wave's tiles first + i*tstride, i < mine: full rounds of the ring
    // in a loop with no early exit (so every path into its head has issued
    // the same loads), then the partial last round
    const uint64_t mine = first < tiles ? (tiles - first + tstride - 1) / tstride : 0;
    const uint64_t rounds = mine / kSlicePrefetch;
    const uint32_t rest = (uint32_t)(mine % kSlicePrefetch);
    auto add = [&](uint64_t tile, f4 v, uint32_t L) {
        const bool valid = tile * 64 + lane < n && L < k;
        if (valid) {
#pragma unroll
            for (uint32_t j = 0; j < kSliceF; ++j)
                if (j < nf) {
                    bool ok;
                    const int32_t q = grid_q(v[j], sc[j], ok);
                    lds_add64(h.acc + j * k + L, (long long)q);
                    if (__builtin_expect(!ok, 0))
                        lds_or(h.flg + j * k + L, special_flag(v[j]));
                }
        }
    };
#pragma unroll
    for (uint32_t s = 0; s < kSlicePrefetch; ++s)
        load(first + s * tstride, vr[s], Lr[s]);
    uint64_t base = first;
    for (uint64_t r = 0; r < rounds; ++r, base += kSlicePrefetch * tstride) {
#pragma unroll
        for (uint32_t s = 0; s < kSlicePrefetch; ++s) {
            const uint64_t tile = base + s * tstride;  // wave-uniform, < tiles
            add(tile, vr[s], Lr[s]);
            // refill the slot once its tile is added: its registers are free
            load(tile + kSlicePrefetch * tstride, vr[s], Lr[s]);
        }
    }
#pragma unroll
    for (uint32_t s = 0; s < kSlicePrefetch; ++s)
        if (s < rest)
            add(base + s * tstride, vr[s], Lr[s]);
    __syncthreads();
    for (uint32_t i = tid; i < k * nf; i += blockDim.x) {
        const uint32_t j = i / k, c = i % k;
        part[(uint64_t)blockIdx.x * ((uint64_t)d * k) + (uint64_t)(f0 + j) * k + c] =
            cell_value(h.acc[i], grid_unit(h.scl[j]), h.flg[i]);
    }
}

using CmKernel = void (*)(const float *, uint64_t, uint32_t, uint32_t, const uint32_t *, GridRef, double *);

// Points per work-item at most (VEC x DMAX values in registers, no spills).
constexpr int cm_vec_cap(int dmax) { return dmax >= 64 ? 1 : dmax >= 32 ? 2 : 4; }

template <int DMAX, int BLOCK>
CmKernel pick_cm_block(int vec)
{
    constexpr int cap = cm_vec_cap(DMAX);
    if constexpr (cap >= 4)
        if (vec >= 4)
            return cu_cluster_merge_kernel<DMAX, 4, BLOCK>;
    if constexpr (cap >= 2)
        if (vec >= 2)
            return cu_cluster_merge_kernel<DMAX, 2, BLOCK>;
    return cu_cluster_merge_kernel<DMAX, 1, BLOCK>;
}

template <int DMAX>
CmKernel pick_cm_vec(int vec, bool b256)
{
    return b256 ? pick_cm_block<DMAX, 256>(vec) : pick_cm_block<DMAX, 0>(vec);
}

CmKernel pick_cm(int dmax, int vec, bool b256)
{
    switch (dmax) {
    case 1: return pick_cm_vec<1>(vec, b256);
    case 2: return pick_cm_vec<2>(vec, b256);
    case 4: return pick_cm_vec<4>(vec, b256);
    case 8: return pick_cm_vec<8>(vec, b256);
    case 16: return pick_cm_vec<16>(vec, b256);
    case 32: return pick_cm_vec<32>(vec, b256);
    case 64: return pick_cm_vec<64>(vec, b256);
    default: return nullptr;
    }
}

}  // namespace

int run_centroid_update(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                        uint64_t k, const float *points, double *sums, const uint32_t *labels,
                        hipStream_t stream)
{
    CLK_REQUIRE(d >= 1 && k >= 1 && k <= 0xFFFFFFFFull && d <= 0xFFFFull, CLK_ERR_SIZE,
                "centroid_update: bad d or k");
    CLK_REQUIRE(points && sums && labels, CLK_ERR_INVALID_ARGUMENT, "centroid_update: null buffer");
    if (n == 0)
        return CLK_OK;
    const int strategy = cfg ? cfg->strategy : CLK_CENTROID_CLUSTER_MERGE;
    CLK_REQUIRE(strategy >= CLK_CENTROID_FEATURE_SUM && strategy <= CLK_CENTROID_CLUSTER_MERGE,
                CLK_ERR_INVALID_ARGUMENT, "centroid_update: unknown strategy id");
    const uint64_t B = (cfg && cfg->local_size[0]) ? cfg->local_size[0] : 256;
    CLK_REQUIRE(B <= (uint64_t)ctx->max_threads_per_block, CLK_ERR_INVALID_ARGUMENT,
                "centroid_update: local_size exceeds the device limit");
    const uint64_t user_grid = (cfg && cfg->global_size[0]) ? div_up(cfg->global_size[0], B) : 0;
    const uint64_t dk = d * k;
    const int dmax = dmax_bucket(d);

    GridRef sg;
    CLK_TRY(resolve_grid(ctx, d, n, points, stream, &sg));
    const uint64_t head = (uint64_t)acc_head_words((uint32_t)d, (uint32_t)k) * 4;

    const void *kern = nullptr;
    const char *name = nullptr;
    size_t lds = 0;
    dim3 grid, block;
    uint32_t tf = 1;
    uint64_t gx = 0;

    if (strategy == CLK_CENTROID_CLUSTER_MERGE && dmax && dk < (1ull << 24) &&
        head + dk * B * 4 <= ctx->lds_per_block) {
        const int vec = std::min(pick_vector_width(cfg ? cfg->vector_length : 4), cm_vec_cap(dmax));
        CmKernel kc = pick_cm(dmax, vec, B == 256);
        lds = head + dk * B * 4;
        gx = user_grid ? user_grid : auto_grid(ctx, (const void *)kc, (int)B, lds, n / vec);
        gx = std::max<uint64_t>(1, std::min<uint64_t>(gx, 1u << 20));
        if (lds > 65536)
            CLK_HIP_CHECK(hipFuncSetAttribute((const void *)kc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        CLK_TRY(ensure_scratch(ctx->part_sums, gx * dk * sizeof(double)));
        double *ps = static_cast<double *>(ctx->part_sums.ptr);
        {
            KernelTimer t(ctx, "CentroidUpdateClusterMerge", stream);
            hipLaunchKernelGGL(kc, dim3((unsigned)gx), dim3((unsigned)B), lds, stream, points, n,
                               (uint32_t)d, (uint32_t)k, labels, sg, ps);
            CLK_HIP_CHECK(hipGetLastError());
        }
        return launch_reduce_f64(ctx, ps, gx, dk, sums, true, stream);
    }
    if (strategy == CLK_CENTROID_FEATURE_SUM_PARDIM) {
        uint64_t lf = (cfg && cfg->local_features) ? cfg->local_features : 1;
        tf = (uint32_t)((cfg && cfg->thread_features) ? cfg->thread_features : 1);
        lf = std::min<uint64_t>(lf, d);
        uint64_t bx = std::max<uint64_t>(1, B / lf);
        uint64_t fs = lf * tf;
        lds = (uint64_t)acc_head_words((uint32_t)fs, (uint32_t)k) * 4 + k * fs * bx * 4;
        if (k * fs < (1ull << 24) && lds <= ctx->lds_per_block && bx * lf <= (uint64_t)ctx->max_threads_per_block) {
            kern = (const void *)cu_pardim_kernel;
            block = dim3((unsigned)bx, (unsigned)lf);
            uint64_t gy = div_up(d, fs);
            gx = user_grid ? user_grid : auto_grid(ctx, kern, (int)(bx * lf), lds, n);
            gx = std::max<uint64_t>(1, std::min<uint64_t>(div_up(gx, gy), 1u << 20));
            grid = dim3((unsigned)gx, (unsigned)gy);
            name = "CentroidUpdateFeatureSumPardim";
        }
    }
    if (!kern && (strategy == CLK_CENTROID_FEATURE_SUM || strategy == CLK_CENTROID_CLUSTER_MERGE ||
                  strategy == CLK_CENTROID_FEATURE_SUM_PARDIM) &&
        dk < (1ull << 24) && head <= ctx->lds_per_block) {
        kern = (const void *)cu_feature_sum_kernel;
        lds = head;
        block = dim3((unsigned)B);
        gx = user_grid ? user_grid : auto_grid(ctx, kern, (int)B, lds, n * d);
        gx = std::max<uint64_t>(1, std::min<uint64_t>(gx, 1u << 20));
        grid = dim3((unsigned)gx);
        name = "CentroidUpdateFeatureSum";
    }
    if (!kern && slice_lds_bytes(k) <= ctx->lds_per_block) {
        // Large k*d: cluster-major feature slices, every strategy.
        const size_t sl = slice_lds_bytes(k);
        if (sl > 65536)
            CLK_HIP_CHECK(hipFuncSetAttribute((const void *)cu_slice_kernel,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sl));
        const uint64_t gy = div_up(d, (uint64_t)kSliceF);
        // at least four work-groups per CU over the slices, a quarter
        // k-th of the points per work-group (partial rows stay few)
        uint64_t gxs = std::max<uint64_t>(div_up(n, 64 * k * kSliceWaves), div_up(4 * (uint64_t)ctx->num_cus, gy));
        gxs = std::max<uint64_t>(1, std::min<uint64_t>(gxs, div_up(n, 64 * kSliceWaves)));
        gxs = std::min<uint64_t>(gxs, 65535);
        CLK_TRY(ensure_scratch(ctx->part_sums, gxs * dk * sizeof(double)));
        double *ps = static_cast<double *>(ctx->part_sums.ptr);
        {
            KernelTimer t(ctx, "CentroidUpdateSlices", stream);
            hipLaunchKernelGGL(cu_slice_kernel, dim3((unsigned)gxs, (unsigned)gy), dim3(64 * kSliceWaves), sl,
                               stream, points, n, (uint32_t)d, (uint32_t)k, labels, sg, ps);
            CLK_HIP_CHECK(hipGetLastError());
        }
        return launch_reduce_f64(ctx, ps, gxs, dk, sums, true, stream);
    }
    if (!kern)
        return launch_accumulate_global(ctx, d, n, k, points, labels, sg, sums, nullptr, stream);

    if (lds > 65536)
        CLK_HIP_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    CLK_TRY(ensure_scratch(ctx->part_sums, gx * dk * sizeof(double)));
    double *ps = static_cast<double *>(ctx->part_sums.ptr);
    {
        KernelTimer t(ctx, name, stream);
        if (kern == (const void *)cu_pardim_kernel)
            hipLaunchKernelGGL(cu_pardim_kernel, grid, block, lds, stream, points, n, (uint32_t)d,
                               (uint32_t)k, tf, labels, sg, ps);
        else
            hipLaunchKernelGGL(cu_feature_sum_kernel, grid, block, lds, stream, points, n,
                               (uint32_t)d, (uint32_t)k, labels, sg, ps);
        CLK_HIP_CHECK(hipGetLastError());
    }
    return launch_reduce_f64(ctx, ps, gx, dk, sums, true, stream);
}

}  // namespace clk

extern "C" int clk_centroid_update(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                                   uint64_t n, uint64_t k, const float *points, double *sums,
                                   const uint32_t *labels, const uint32_t *masses, void *stream)
{
    (void)masses;
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(!cfg || cfg->stage == CLK_STAGE_CENTROID_UPDATE, CLK_ERR_INVALID_ARGUMENT,
                "clk_centroid_update: config is not a centroid-update-stage config");
    return clk::run_centroid_update(ctx, cfg, d, n, k, points, sums, labels,
                                    clk::as_stream(stream));
}
