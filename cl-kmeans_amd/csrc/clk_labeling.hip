// Labeling pass (LabelingUnrollVector / lloyd_labeling_vp_clcp.cl) and the
// global-accumulation pass used when k*d exceeds the LDS tiles.
#include "clk_common.hpp"
#include "clk_grid.hpp"
#include "clk_kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace clk {
namespace {

// Centroids staged in LDS ([k][DMAX]); each work-item labels VEC
// consecutive points per step with 128-bit row loads.  VEC * DMAX point
// values live in registers: with 64 features the work-group is capped at 512
// work-items so that the kernel may use 256 registers.
template <int DMAX>
constexpr int label_lds_threads() { return DMAX >= 64 ? 512 : 1024; }

// BOUNDED ("unroll_vector_bounded"): the labels buffer is read first; a
// point whose distance to its previous label's centroid is within that
// centroid's keep threshold (keep_threshold: the scan provably returns the
// same label) is not scanned, and only changed labels are stored.
template <int DMAX, int VEC, bool BOUNDED>
__global__ void __launch_bounds__(label_lds_threads<DMAX>())
label_lds_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                 const float *__restrict__ cen, uint32_t *__restrict__ labels)
{
    extern __shared__ __align__(16) float cen_lds[];
    float *thr = cen_lds + ((k * DMAX + 3u) & ~3u);  // BOUNDED: [k]
    stage_centroids<DMAX>(cen, d, k, cen_lds);
    __syncthreads();
    if constexpr (BOUNDED) {
        stage_keep_thresholds<DMAX>(cen_lds, k, thr);
        __syncthreads();
    }
    // V points p .. p+V-1 (V = VEC, or 1 for the n % VEC tail points)
    auto label_points = [&](auto vtag, uint64_t p) {
        constexpr int V = decltype(vtag)::value;
        float x[DMAX][V];
        load_points<DMAX, V>(pts, n, d, p, x);
        uint32_t lab[V];
        if constexpr (BOUNDED) {
            uint32_t prev[V];
            stream_load_labels<V>(labels, p, prev);
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
                keep = keep && valid && dist <= thr[a];
                lab[v] = a;
            }
            if (!keep)
                argmin_lds<DMAX, V>(x, cen_lds, k, lab);
            bool changed = false;
#pragma unroll
            for (int v = 0; v < V; ++v)
                changed = changed || lab[v] != prev[v];
            if (changed)
                store_labels<V>(labels, p, lab);
        } else {
            argmin_lds<DMAX, V>(x, cen_lds, k, lab);
            store_labels<V>(labels, p, lab);
        }
    };
    const uint64_t groups = n / VEC;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t g = id; g < groups; g += stride)
        label_points(std::integral_constant<int, VEC>{}, g * VEC);
    if (VEC > 1 && id == groups % stride)
        for (uint64_t p = groups * VEC; p < n; ++p)
            label_points(std::integral_constant<int, 1>{}, p);
}

// k*d beyond LDS: the centroids pass through LDS in chunks of `chunk`
// (feature-padded rows [chunk][DMAX]) while each work-item keeps its VEC
// points and their running minima in registers.  The loop over point groups
// is work-group uniform (every work-item takes part in every staging).
template <int DMAX, int VEC>
__global__ void __launch_bounds__(256)
label_chunked_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                     const float *__restrict__ cen, uint32_t *__restrict__ labels, uint32_t chunk)
{
    extern __shared__ __align__(16) float cen_lds[];
    const uint64_t per = (uint64_t)blockDim.x * VEC;
    for (uint64_t base = (uint64_t)blockIdx.x * per; base < n; base += (uint64_t)gridDim.x * per) {
        const uint64_t p = base + (uint64_t)threadIdx.x * VEC;
        const bool full = p + VEC <= n;  // else the last, partial group or none
        float x[DMAX][VEC];
        if (full) {
            load_points<DMAX, VEC>(pts, n, d, p, x);
        } else {
#pragma unroll
            for (int f = 0; f < DMAX; ++f)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                    x[f][v] = (f < (int)d && p + v < n) ? pts[(uint64_t)f * n + p + v] : 0.0f;
        }
        float best[VEC];
        uint32_t lab[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            best[v] = FLT_MAX;
            lab[v] = 0;
        }
        for (uint32_t c0 = 0; c0 < k; c0 += chunk) {
            const uint32_t cn = min(chunk, k - c0);
            __syncthreads();  // the previous chunk has been read
            for (uint32_t i = threadIdx.x; i < cn * DMAX; i += blockDim.x) {
                const uint32_t c = i / DMAX, f = i % DMAX;
                cen_lds[i] = f < d ? cen[(uint64_t)f * k + c0 + c] : 0.0f;
            }
            __syncthreads();
            argmin_lds_chunk<DMAX, VEC>(x, cen_lds, c0, cn, best, lab);
        }
        if (full) {
            store_labels<VEC>(labels, p, lab);
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                if (p + v < n)
                    labels[p + v] = lab[v];
        }
    }
}

// Centroids read from global memory with work-group-uniform addresses (any
// k).  DMAX > 0 keeps the point in registers; DMAX == 0 re-reads it (d > 64).
template <int DMAX>
__global__ void __launch_bounds__(256)
label_global_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                    const float *__restrict__ cen, uint32_t *__restrict__ labels)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        float best = FLT_MAX;
        uint32_t lab = 0;
        if constexpr (DMAX > 0) {
            float x[DMAX];
#pragma unroll
            for (int f = 0; f < DMAX; ++f)
                x[f] = (f < (int)d) ? pts[(uint64_t)f * n + p] : 0.0f;
            for (uint32_t c = 0; c < k; ++c) {
                float t = x[0] - cen[c];
                float dist = t * t;
#pragma unroll
                for (int f = 1; f < DMAX; ++f) {
                    if (f < (int)d) {
                        t = x[f] - cen[(uint64_t)f * k + c];
                        dist = __builtin_fmaf(t, t, dist);
                    }
                }
                bool lt = dist < best;
                best = lt ? dist : best;
                lab = lt ? c : lab;
            }
        } else {
            for (uint32_t c = 0; c < k; ++c) {
                float t = pts[p] - cen[c];
                float dist = t * t;
                for (uint32_t f = 1; f < d; ++f) {
                    t = pts[(uint64_t)f * n + p] - cen[(uint64_t)f * k + c];
                    dist = __builtin_fmaf(t, t, dist);
                }
                bool lt = dist < best;
                best = lt ? dist : best;
                lab = lt ? c : lab;
            }
        }
        labels[p] = lab;
    }
}

__global__ void __launch_bounds__(256)
accumulate_global_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                         const uint32_t *__restrict__ labels, GridRef grid, double *__restrict__ sums,
                         uint32_t *__restrict__ masses)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const uint32_t c = labels[p];
        if (c >= k)
            continue;
        if (masses)
            atomicAdd(&masses[c], 1u);
        for (uint32_t f = 0; f < d; ++f) {
            // the grid value (a multiple of the unit below 2^53 units: every
            // partial sum is exact); off the grid the special value
            const float x = pts[(uint64_t)f * n + p];
            const float sc = grid_scale(grid.gmax[f], grid.bits);
            bool ok;
            const int32_t q = grid_q(x, sc, ok);
            const double v = ok ? (double)q * grid_unit(sc) : special_value(special_flag(x));
            unsafeAtomicAdd(&sums[(uint64_t)f * k + c], v);
        }
    }
}

using LabelKernel = void (*)(const float *, uint64_t, uint32_t, uint32_t, const float *,
                             uint32_t *);

// Points per work-item at most (VEC x DMAX values in registers, no spills).
constexpr int label_vec_cap(int dmax) { return dmax >= 64 ? 1 : dmax >= 32 ? 2 : 4; }

template <int DMAX, bool BOUNDED>
LabelKernel pick_label_vec(int vec)
{
    constexpr int cap = label_vec_cap(DMAX);
    if constexpr (cap >= 4)
        if (vec >= 4)
            return label_lds_kernel<DMAX, 4, BOUNDED>;
    if constexpr (cap >= 2)
        if (vec >= 2)
            return label_lds_kernel<DMAX, 2, BOUNDED>;
    return label_lds_kernel<DMAX, 1, BOUNDED>;
}

template <bool BOUNDED>
LabelKernel pick_label_lds(int dmax, int vec)
{
    switch (dmax) {
    case 1: return pick_label_vec<1, BOUNDED>(vec);
    case 2: return pick_label_vec<2, BOUNDED>(vec);
    case 4: return pick_label_vec<4, BOUNDED>(vec);
    case 8: return pick_label_vec<8, BOUNDED>(vec);
    case 16: return pick_label_vec<16, BOUNDED>(vec);
    case 32: return pick_label_vec<32, BOUNDED>(vec);
    case 64: return pick_label_vec<64, BOUNDED>(vec);
    default: return nullptr;
    }
}

using ChunkKernel = void (*)(const float *, uint64_t, uint32_t, uint32_t, const float *, uint32_t *,
                             uint32_t);

template <int DMAX>
ChunkKernel pick_chunk_vec(int vec)
{
    constexpr int cap = label_vec_cap(DMAX);
    if constexpr (cap >= 4)
        if (vec >= 4)
            return label_chunked_kernel<DMAX, 4>;
    if constexpr (cap >= 2)
        if (vec >= 2)
            return label_chunked_kernel<DMAX, 2>;
    return label_chunked_kernel<DMAX, 1>;
}

ChunkKernel pick_label_chunked(int dmax, int vec)
{
    switch (dmax) {
    case 1: return pick_chunk_vec<1>(vec);
    case 2: return pick_chunk_vec<2>(vec);
    case 4: return pick_chunk_vec<4>(vec);
    case 8: return pick_chunk_vec<8>(vec);
    case 16: return pick_chunk_vec<16>(vec);
    case 32: return pick_chunk_vec<32>(vec);
    case 64: return pick_chunk_vec<64>(vec);
    default: return nullptr;
    }
}

LabelKernel pick_label_global(int dmax)
{
    switch (dmax) {
    case 1: return label_global_kernel<1>;
    case 2: return label_global_kernel<2>;
    case 4: return label_global_kernel<4>;
    case 8: return label_global_kernel<8>;
    case 16: return label_global_kernel<16>;
    case 32: return label_global_kernel<32>;
    case 64: return label_global_kernel<64>;
    default: return label_global_kernel<0>;
    }
}

}  // namespace

int launch_labeling(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                    uint64_t k, const float *points, const float *centroids, uint32_t *labels,
                    hipStream_t stream, bool bounded)
{
    if (n == 0)
        return CLK_OK;
    const uint64_t B = (cfg && cfg->local_size[0]) ? cfg->local_size[0] : 256;
    CLK_REQUIRE(B <= (uint64_t)ctx->max_threads_per_block, CLK_ERR_INVALID_ARGUMENT,
                "labeling: local_size exceeds the device limit");
    const int dmax = dmax_bucket(d);
    const uint64_t cen_plain = dmax ? ((k * dmax + 3) & ~3ull) * 4 : 0;
    // the bounded variant adds k keep thresholds; beyond LDS it runs as the
    // exhaustive kernels below (the same labels)
    bounded = bounded && dmax && cen_plain + k * 4 <= ctx->lds_per_block;
    const uint64_t cen_bytes = cen_plain + (bounded ? k * 4 : 0);
    if (dmax && cen_bytes <= ctx->lds_per_block) {
        const int vec = std::min(pick_vector_width(cfg ? cfg->vector_length : 4),
                                 label_vec_cap(dmax));
        LabelKernel kern = bounded ? pick_label_lds<true>(dmax, vec) : pick_label_lds<false>(dmax, vec);
        CLK_REQUIRE(dmax < 64 || B <= 512, CLK_ERR_INVALID_ARGUMENT,
                    "labeling: local_size above 512 with more than 32 features");
        uint64_t grid = (cfg && cfg->global_size[0])
                            ? div_up(cfg->global_size[0], B)
                            : auto_grid(ctx, (const void *)kern, (int)B, cen_bytes, n / vec);
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, 1u << 20));
        KernelTimer t(ctx, bounded ? "LabelingUnrollVectorBounded" : "LabelingUnrollVector", stream);
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3((unsigned)B), cen_bytes, stream,
                           points, n, (uint32_t)d, (uint32_t)k, centroids, labels);
        CLK_HIP_CHECK(hipGetLastError());
        return CLK_OK;
    }
    if (dmax) {  // centroid chunks through LDS (64 KiB: two work-groups per CU)
        const int vec = std::min(pick_vector_width(cfg ? cfg->vector_length : 4),
                                 label_vec_cap(dmax));
        ChunkKernel kern = pick_label_chunked(dmax, vec);
        const uint32_t chunk = (uint32_t)std::min<uint64_t>(k, (64u << 10) / (4u * dmax));
        const size_t lds = (size_t)chunk * dmax * 4;
        uint64_t grid = auto_grid(ctx, (const void *)kern, 256, lds, n / vec);
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, 1u << 20));
        KernelTimer t(ctx, "LabelingChunked", stream);
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), lds, stream, points, n, (uint32_t)d,
                           (uint32_t)k, centroids, labels, chunk);
        CLK_HIP_CHECK(hipGetLastError());
        return CLK_OK;
    }
    LabelKernel kern = pick_label_global(dmax);
    const uint64_t Bg = 256;
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(div_up(n, Bg), (uint64_t)ctx->num_cus * 8));
    KernelTimer t(ctx, "LabelingGlobal", stream);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3((unsigned)Bg), 0, stream, points, n,
                       (uint32_t)d, (uint32_t)k, centroids, labels);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

int launch_accumulate_global(clk_context *ctx, uint64_t d, uint64_t n, uint64_t k,
                             const float *points, const uint32_t *labels, const GridRef &sgrid,
                             double *sums, uint32_t *masses, hipStream_t stream)
{
    if (n == 0)
        return CLK_OK;
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(div_up(n, 256), (uint64_t)ctx->num_cus * 8));
    KernelTimer t(ctx, "AccumulateGlobal", stream);
    hipLaunchKernelGGL(accumulate_global_kernel, dim3((unsigned)grid), dim3(256), 0, stream,
                       points, n, (uint32_t)d, (uint32_t)k, labels, sgrid, sums, masses);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

}  // namespace clk

extern "C" int clk_labeling(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                            uint64_t n, uint64_t k, const float *points,
                            const float *centroids, uint32_t *labels, void *stream)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(!cfg || cfg->stage == CLK_STAGE_LABELING, CLK_ERR_INVALID_ARGUMENT,
                "clk_labeling: config is not a labeling-stage config");
    const int strategy = cfg ? cfg->strategy : CLK_LABELING_UNROLL_VECTOR;
    CLK_REQUIRE(strategy == CLK_LABELING_UNROLL_VECTOR || strategy == CLK_LABELING_MFMA_BF16 ||
                    strategy == CLK_LABELING_MFMA_FP16 || strategy == CLK_LABELING_UNROLL_VECTOR_BOUNDED,
                CLK_ERR_INVALID_ARGUMENT, "clk_labeling: unknown strategy id");
    CLK_REQUIRE(d >= 1 && k >= 1 && k <= 0xFFFFFFFFull, CLK_ERR_SIZE, "labeling: bad d or k");
    CLK_REQUIRE(points && centroids && labels, CLK_ERR_INVALID_ARGUMENT, "labeling: null buffer");
    const bool mfma = strategy == CLK_LABELING_MFMA_BF16 || strategy == CLK_LABELING_MFMA_FP16;
    if (mfma && clk::mfma_labeling_supported(ctx, d, n, k))
        return clk::launch_labeling_mfma(ctx, strategy == CLK_LABELING_MFMA_FP16, d, n, k, points,
                                         centroids, labels, clk::as_stream(stream));
    return clk::launch_labeling(ctx, cfg, d, n, k, points, centroids, labels, clk::as_stream(stream),
                                strategy == CLK_LABELING_UNROLL_VECTOR_BOUNDED);
}
