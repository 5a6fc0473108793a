// Centroid update: sums[f,c] += sum_{p : labels[p] == c} x_f, the three
// strategies of CentroidUpdateFactory built for gfx950:
//   cluster_merge       one LDS tile column [k*d] per work-item (no
//                       collisions), all features of the work-item's points
//                       (lloyd_cluster_merge.cl's private k x d tile)
//   feature_sum_pardim  2-D grid: blockIdx.y selects a slice of
//                       local_features*thread_features features, threadIdx.y
//                       a thread_features sub-slice; private LDS columns
//                       (lloyd_feature_sum_pardim.cl's points x feature tiles)
//   feature_sum         one LDS tile [d][k] shared by the work-group, work-
//                       items stride over the flattened feature-major matrix
//                       (coalesced rows), ds_add_f32 resolves collisions; the
//                       smallest LDS footprint, for large k*d
// Tiles are summed per work-group in FP64 and merged by a fixed-order column
// reduction, except feature_sum whose in-tile FP32 addition order is the
// hardware's.  For large k*d (when neither fits) every strategy uses the
// cluster-major slice kernel below (per-wave LDS tables [k][4] per feature
// slice, updated by plain read-modify-writes made collision-free by an owner
// claim; LDS float atomics run ~9x slower than read-modify-writes on
// gfx950); the FP64 global-atomic pass remains for k too large for that.
// `masses` is not read, as in the reference kernels.
#include "clk_common.hpp"
#include "clk_kernels.hpp"

#include <algorithm>

namespace clk {
namespace {

template <int DMAX, int VEC, int BLOCK>
__global__ void __launch_bounds__(BLOCK > 0 ? BLOCK : 1024)
cu_cluster_merge_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                        const uint32_t *__restrict__ labels, double *__restrict__ part)
{
    extern __shared__ __align__(16) float tile[];  // [k*d][B]
    const uint32_t B = BLOCK > 0 ? BLOCK : blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t rows = k * d;
    for (uint32_t i = tid; i < rows * B; i += B)
        tile[i] = 0.0f;
    __syncthreads();
    float *mine = tile + tid;
    const uint32_t row_stride = d * B;
    const uint64_t groups = n / VEC;
    const uint64_t stride = (uint64_t)gridDim.x * B;
    for (uint64_t g = (uint64_t)blockIdx.x * B + tid; g < groups; g += stride) {
        const uint64_t p = g * VEC;
        float x[DMAX][VEC];
        load_points<DMAX, VEC>(pts, n, d, p, x);
        uint32_t lab[VEC];
        load_labels<VEC>(labels, p, lab);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            if (lab[v] >= k)
                continue;
            float *t = mine + __umul24(lab[v], row_stride);
#pragma unroll
            for (int f = 0; f < DMAX; ++f)
                if (f < (int)d)
                    lds_acc(t + f * B, x[f][v]);
        }
    }
    __syncthreads();
    for (uint32_t r = tid; r < rows; r += B) {
        const uint32_t c = r / d, f = r % d;
        part[(uint64_t)blockIdx.x * rows + (uint64_t)f * k + c] =
            lds_row_sum_f32(tile + (uint64_t)r * B, B, r);
    }
}

// blockDim = (Bx, lf); block (x, y) owns features [y*FS, y*FS+FS), FS = lf*tf;
// thread (tx, ty) owns features y*FS + ty*tf + [0, tf) of points
// x*Bx + tx + i*gridDim.x*Bx.  LDS rows c*FS + (ty*tf + j), columns tx.
__global__ void __launch_bounds__(1024)
cu_pardim_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k, uint32_t tf,
                 const uint32_t *__restrict__ labels, double *__restrict__ part)
{
    extern __shared__ __align__(16) float tile[];
    const uint32_t Bx = blockDim.x, lf = blockDim.y;
    const uint32_t FS = lf * tf;
    const uint32_t f0 = blockIdx.y * FS;
    const uint32_t tlin = threadIdx.y * Bx + threadIdx.x;
    const uint32_t nthreads = Bx * lf;
    const uint32_t rows = k * FS;
    for (uint32_t i = tlin; i < rows * Bx; i += nthreads)
        tile[i] = 0.0f;
    __syncthreads();
    const uint32_t fsub = threadIdx.y * tf;  // first feature of this thread in the slice
    const uint64_t stride = (uint64_t)gridDim.x * Bx;
    for (uint64_t p = (uint64_t)blockIdx.x * Bx + threadIdx.x; p < n; p += stride) {
        const uint32_t c = labels[p];
        if (c >= k)
            continue;
        float *t = tile + ((uint64_t)c * FS + fsub) * Bx + threadIdx.x;
        for (uint32_t j = 0; j < tf; ++j) {
            const uint32_t f = f0 + fsub + j;
            if (f < d)
                lds_acc(t + j * Bx, pts[(uint64_t)f * n + p]);
        }
    }
    __syncthreads();
    for (uint32_t r = tlin; r < rows; r += nthreads) {
        const uint32_t c = r / FS, f = f0 + r % FS;
        if (f < d)
            part[(uint64_t)blockIdx.x * ((uint64_t)d * k) + (uint64_t)f * k + c] =
                lds_row_sum_f32(tile + (uint64_t)r * Bx, Bx, r);
    }
}

// Shared [d][k] tile; work-items stride over elements e = f*n + p.
__global__ void __launch_bounds__(1024)
cu_feature_sum_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                      const uint32_t *__restrict__ labels, double *__restrict__ part)
{
    extern __shared__ __align__(16) float tile[];
    const uint32_t rows = d * k;
    for (uint32_t i = threadIdx.x; i < rows; i += blockDim.x)
        tile[i] = 0.0f;
    __syncthreads();
    const uint64_t total = n * d;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint64_t f = e / n, p = e - f * n;
        const uint32_t c = labels[p];
        if (c < k)
            lds_add(tile + f * k + c, pts[e]);
    }
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < rows; r += blockDim.x)
        part[(uint64_t)blockIdx.x * rows + r] = (double)tile[r];
}

// Cluster-major feature slices for large k*d.  blockIdx.y selects features
// [4*blockIdx.y, +4); each wave owns a private FP32 table T[c] (one float4 per
// cluster) and an owner array O[c] in LDS, so no two waves ever touch one
// table.  A wave takes 64 consecutive points (lane = point: 4 coalesced
// feature loads and the label) and adds each point's float4 into T[label]
// with a plain read-modify-write: lanes write their id to O[label], read it
// back, and the lanes that find their own id -- one per distinct label --
// update their row; the others retry (at most the largest label
// multiplicity in the wave, usually 1-3 rounds).  The wave tables are summed
// in FP64 into the work-group's partial row.  Work-groups are sized so a
// table cell takes about 64 points (FP32 partial sums with few roundings).
constexpr uint32_t kSliceF = 4, kSliceWaves = 4, kSlicePrefetch = 8;

size_t slice_lds_bytes(uint64_t k)
{
    return (size_t)kSliceWaves * k * (kSliceF * 4 + 4);
}

__global__ void __launch_bounds__(64 * kSliceWaves)
cu_slice_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                const uint32_t *__restrict__ labels, double *__restrict__ part)
{
    extern __shared__ __align__(16) float tab[];  // [waves][k] float4, then [waves][k] owners
    typedef float f4 __attribute__((ext_vector_type(4)));
    // the wave index as a scalar: tile indices and the loop below are then
    // wave-uniform, so the prefetch ring needs no divergent control flow
    const uint32_t tid = threadIdx.x, lane = tid % 64;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid / 64));
    f4 *T = reinterpret_cast<f4 *>(tab) + (size_t)wave * k;
    uint32_t *O = reinterpret_cast<uint32_t *>(reinterpret_cast<f4 *>(tab) + (size_t)kSliceWaves * k) +
                  (size_t)wave * k;
    for (uint32_t i = tid; i < kSliceWaves * k; i += blockDim.x)
        reinterpret_cast<f4 *>(tab)[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();
    const uint32_t f0 = blockIdx.y * kSliceF;
    const uint32_t nf = min(kSliceF, d - f0);
    const uint64_t tiles = (n + 63) / 64;
    const uint64_t tstride = (uint64_t)gridDim.x * kSliceWaves;
    // Every load is unconditional and in bounds (a tile past the end reads
    // the last tile, a feature past the slice reads its last feature; such
    // values are never added or are zeroed at use), so no branch separates
    // a load from its use and the loads of later tiles stay in flight.
    auto load = [&](uint64_t tile, f4 &v, uint32_t &L) {
        const uint64_t p = min(min(tile, tiles - 1) * 64 + lane, n - 1);
#pragma unroll
        for (uint32_t j = 0; j < kSliceF; ++j)
            v[j] = pts[(uint64_t)(f0 + min(j, nf - 1)) * n + p];
        L = labels[p];
    };
    // A ring of kSlicePrefetch tiles in registers: the loads of the next
    // kSlicePrefetch tiles of this wave are in flight while one is added
    // (one tile alone leaves too few bytes in flight to cover HBM latency
    // at the few waves per CU the tables allow).  The ring is indexed by
    // the unrolled slot, so it stays in registers.
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
#pragma unroll
        for (uint32_t j = 0; j < kSliceF; ++j)
            v[j] = j < nf ? v[j] : 0.0f;
        const bool valid = tile * 64 + lane < n;
        bool pending = valid && L < k;
        while (__ballot(pending) != 0) {
            if (pending)
                O[L] = lane;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const bool win = pending && O[L] == lane;
            if (win)
                T[L] = T[L] + v;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            pending = pending && !win;
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
    const f4 *all = reinterpret_cast<const f4 *>(tab);
    for (uint32_t i = tid; i < k * nf; i += blockDim.x) {
        const uint32_t c = i / nf, j = i % nf;
        double s = 0.0;
        for (uint32_t w = 0; w < kSliceWaves; ++w)
            s += (double)all[(size_t)w * k + c][j];
        part[(uint64_t)blockIdx.x * ((uint64_t)d * k) + (uint64_t)(f0 + j) * k + c] = s;
    }
}

using CmKernel = void (*)(const float *, uint64_t, uint32_t, uint32_t, const uint32_t *, double *);

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

    const void *kern = nullptr;
    const char *name = nullptr;
    size_t lds = 0;
    dim3 grid, block;
    uint32_t tf = 1;
    uint64_t gx = 0;

    if (strategy == CLK_CENTROID_CLUSTER_MERGE && dmax && dk * B * 4 <= ctx->lds_per_block) {
        const int vec = std::min(pick_vector_width(cfg ? cfg->vector_length : 4, n, points, labels),
                                 cm_vec_cap(dmax));
        CmKernel kc = pick_cm(dmax, vec, B == 256);
        lds = dk * B * 4;
        gx = user_grid ? user_grid : auto_grid(ctx, (const void *)kc, (int)B, lds, n / vec);
        gx = std::max<uint64_t>(1, std::min<uint64_t>(gx, 1u << 20));
        CLK_TRY(ensure_scratch(ctx->part_sums, gx * dk * sizeof(double)));
        double *ps = static_cast<double *>(ctx->part_sums.ptr);
        {
            KernelTimer t(ctx, "CentroidUpdateClusterMerge", stream);
            hipLaunchKernelGGL(kc, dim3((unsigned)gx), dim3((unsigned)B), lds, stream, points, n,
                               (uint32_t)d, (uint32_t)k, labels, ps);
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
        lds = k * fs * bx * 4;
        if (lds <= ctx->lds_per_block && bx * lf <= (uint64_t)ctx->max_threads_per_block) {
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
        dk * 4 <= ctx->lds_per_block) {
        kern = (const void *)cu_feature_sum_kernel;
        lds = dk * 4;
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
        // about 64 points per table cell and wave, at least two resident
        // rounds of work-groups
        uint64_t gxs = std::max<uint64_t>(div_up(n, 64 * k * kSliceWaves), div_up(4 * (uint64_t)ctx->num_cus, gy));
        gxs = std::max<uint64_t>(1, std::min<uint64_t>(gxs, div_up(n, 64 * kSliceWaves)));
        gxs = std::min<uint64_t>(gxs, 65535);
        CLK_TRY(ensure_scratch(ctx->part_sums, gxs * dk * sizeof(double)));
        double *ps = static_cast<double *>(ctx->part_sums.ptr);
        {
            KernelTimer t(ctx, "CentroidUpdateSlices", stream);
            hipLaunchKernelGGL(cu_slice_kernel, dim3((unsigned)gxs, (unsigned)gy), dim3(64 * kSliceWaves), sl,
                               stream, points, n, (uint32_t)d, (uint32_t)k, labels, ps);
            CLK_HIP_CHECK(hipGetLastError());
        }
        return launch_reduce_f64(ctx, ps, gxs, dk, sums, true, stream);
    }
    if (!kern)
        return launch_accumulate_global(ctx, d, n, k, points, labels, sums, nullptr, stream);

    CLK_TRY(ensure_scratch(ctx->part_sums, gx * dk * sizeof(double)));
    double *ps = static_cast<double *>(ctx->part_sums.ptr);
    {
        KernelTimer t(ctx, name, stream);
        if (kern == (const void *)cu_pardim_kernel)
            hipLaunchKernelGGL(cu_pardim_kernel, grid, block, lds, stream, points, n, (uint32_t)d,
                               (uint32_t)k, tf, labels, ps);
        else
            hipLaunchKernelGGL(cu_feature_sum_kernel, grid, block, lds, stream, points, n,
                               (uint32_t)d, (uint32_t)k, labels, ps);
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
