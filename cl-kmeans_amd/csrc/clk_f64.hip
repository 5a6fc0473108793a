// The double / uint64 / uint64 type triple of the reference (bench.cpp:469-503
// instantiates every strategy for it; the kernels are the same OpenCL sources
// built with CL_POINT=double, CL_POINT_MAX=DBL_MAX, CL_LABEL=ulong,
// CL_MASS=ulong -- fused_feature_sum.hpp:80-95, labeling_unroll_vector.hpp:
// 64-76).  One streaming pass serves every operator:
//
//   f64_pass_kernel<DMAX, LABEL, SUMS, MASS>
//     LABEL: labels[p] = argmin_c of the sequential fma over f of (x_f-c_f)^2
//            in FP64 from 0, strict '<' against a running minimum from
//            DBL_MAX, label 0 when nothing is smaller (kmeans_naive.cpp:47-63)
//            -- otherwise the labels are read;
//     SUMS:  sums[f][c] += x_f over the points of c;
//     MASS:  masses[c] += #points of c.
//
// Centroids are staged in LDS as [k][DMAX] (zero-padded features add
// fma(0,0,dist) == dist exactly); each work-group accumulates an LDS tile
// (FP64 sums with ds_add_f64, uint32 counts) and writes it as a partial row;
// a fixed-order column reduction adds the rows into the caller's buffers.
// Where the tile does not fit, the pass adds straight into the caller's
// buffers with global FP64 / uint64 atomics (order not fixed).  Labels are
// bit-exact either way; sums differ from a serial sum only by FP64 rounding.
#include "clk_common.hpp"
#include "clk_kernels.hpp"

#include <algorithm>
#include <cfloat>

namespace clk {
namespace {

template <int DMAX, bool LABEL, bool SUMS, bool MASS, bool TILE>
__global__ void __launch_bounds__(1024)
f64_pass_kernel(const double *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                const double *__restrict__ cen, uint64_t *__restrict__ labels,
                double *__restrict__ part_sums, uint32_t *__restrict__ part_masses,
                double *__restrict__ sums, uint64_t *__restrict__ masses)
{
    extern __shared__ __align__(16) double smem64[];
    const uint32_t B = blockDim.x, tid = threadIdx.x;
    double *cen_lds = smem64;                                        // LABEL: [k][DMAX]
    double *tile = cen_lds + (LABEL ? (size_t)k * DMAX : 0);         // SUMS && TILE: [d][k]
    uint32_t *cnt = reinterpret_cast<uint32_t *>(tile + (SUMS && TILE ? (size_t)d * k : 0));  // MASS && TILE

    if constexpr (LABEL)
        for (uint32_t i = tid; i < k * DMAX; i += B) {
            const uint32_t c = i / DMAX, f = i % DMAX;
            cen_lds[i] = f < d ? cen[(size_t)f * k + c] : 0.0;
        }
    if constexpr (TILE) {
        if constexpr (SUMS)
            for (uint32_t i = tid; i < d * k; i += B)
                tile[i] = 0.0;
        if constexpr (MASS)
            for (uint32_t i = tid; i < k; i += B)
                cnt[i] = 0u;
    }
    __syncthreads();

    const uint64_t stride = (uint64_t)gridDim.x * B;
    for (uint64_t p = (uint64_t)blockIdx.x * B + tid; p < n; p += stride) {
        double x[DMAX];
        if constexpr (LABEL || SUMS) {
#pragma unroll
            for (int f = 0; f < DMAX; ++f)
                x[f] = f < (int)d ? stream_load(pts + (uint64_t)f * n + p) : 0.0;
        }
        uint64_t lab;
        if constexpr (LABEL) {
            double best = DBL_MAX;
            uint32_t bl = 0;
            for (uint32_t c = 0; c < k; ++c) {
                const double *cc = cen_lds + (size_t)c * DMAX;
                double dist = 0.0;
#pragma unroll
                for (int f = 0; f < DMAX; ++f) {
                    const double t = x[f] - cc[f];
                    dist = __builtin_fma(t, t, dist);
                }
                if (dist < best) {
                    best = dist;
                    bl = c;
                }
            }
            lab = bl;
            stream_store(lab, labels + p);
        } else {
            lab = labels[p];
            if (lab >= k)  // out-of-range labels are skipped (the reference would write out of bounds)
                continue;
        }
        if constexpr (SUMS) {
#pragma unroll
            for (int f = 0; f < DMAX; ++f)
                if (f < (int)d) {
                    if constexpr (TILE)
                        atomicAdd(&tile[(size_t)f * k + lab], x[f]);
                    else
                        atomicAdd(&sums[(size_t)f * k + lab], x[f]);
                }
        }
        if constexpr (MASS) {
            if constexpr (TILE)
                atomicAdd(&cnt[lab], 1u);
            else
                atomicAdd(reinterpret_cast<unsigned long long *>(&masses[lab]), 1ull);
        }
    }

    if constexpr (TILE && (SUMS || MASS)) {
        __syncthreads();
        if constexpr (SUMS)
            for (uint32_t i = tid; i < d * k; i += B)
                part_sums[(size_t)blockIdx.x * d * k + i] = tile[i];
        if constexpr (MASS)
            for (uint32_t i = tid; i < k; i += B)
                part_masses[(size_t)blockIdx.x * k + i] = cnt[i];
    }
}

// The pass for d beyond the register buckets: centroids and points from
// global memory, global atomics.
template <bool LABEL, bool SUMS, bool MASS>
__global__ void __launch_bounds__(1024)
f64_pass_global_kernel(const double *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                       const double *__restrict__ cen, uint64_t *__restrict__ labels,
                       double *__restrict__ sums, uint64_t *__restrict__ masses)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        uint64_t lab;
        if constexpr (LABEL) {
            double best = DBL_MAX;
            uint32_t bl = 0;
            for (uint32_t c = 0; c < k; ++c) {
                double dist = 0.0;
                for (uint32_t f = 0; f < d; ++f) {
                    const double t = pts[(uint64_t)f * n + p] - cen[(size_t)f * k + c];
                    dist = __builtin_fma(t, t, dist);
                }
                if (dist < best) {
                    best = dist;
                    bl = c;
                }
            }
            lab = bl;
            labels[p] = lab;
        } else {
            lab = labels[p];
            if (lab >= k)
                continue;
        }
        if constexpr (SUMS)
            for (uint32_t f = 0; f < d; ++f)
                atomicAdd(&sums[(size_t)f * k + lab], pts[(uint64_t)f * n + p]);
        if constexpr (MASS)
            atomicAdd(reinterpret_cast<unsigned long long *>(&masses[lab]), 1ull);
    }
}

// masses[c] += sum over rows of part[row * k + c] (fixed order).
__global__ void __launch_bounds__(256)
reduce_masses_u64_kernel(const uint32_t *__restrict__ part, uint64_t rows, uint32_t k,
                         uint64_t *__restrict__ masses)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k)
        return;
    uint64_t s = 0;
    for (uint64_t r = 0; r < rows; ++r)
        s += part[r * k + c];
    masses[c] += s;
}

// centroids[f,c] = sums[f,c] / (double)masses[c] (0/0 = NaN for an empty
// cluster: matrix_binary_op.cl:36-38 with CL_POINT=double); with old, the
// largest ||new_c - old_c|| over non-empty clusters into *max_shift (one
// work-group, fixed order).
__global__ void __launch_bounds__(256)
divide_f64_kernel(const double *__restrict__ sums, const uint64_t *__restrict__ masses, uint32_t d,
                  uint32_t k, double *__restrict__ centroids, const double *__restrict__ old,
                  double *__restrict__ max_shift)
{
    __shared__ double red[256];
    double local = 0.0;
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) {
        const double m = (double)masses[c];
        double sq = 0.0;
        for (uint32_t f = 0; f < d; ++f) {
            const double v = sums[(size_t)f * k + c] / m;
            if (old) {
                const double t = v - old[(size_t)f * k + c];
                sq += t * t;
            }
            centroids[(size_t)f * k + c] = v;
        }
        if (masses[c] != 0)
            local = fmax(local, sqrt(sq));
    }
    if (!max_shift || !old)
        return;
    red[threadIdx.x] = local;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s /= 2) {
        if (threadIdx.x < s)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *max_shift = red[0];
}

__global__ void add_u64_kernel(uint64_t *__restrict__ fst, const uint64_t *__restrict__ snd, uint64_t count)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        fst[i] += snd[i];
}

using PassKernel = void (*)(const double *, uint64_t, uint32_t, uint32_t, const double *, uint64_t *,
                            double *, uint32_t *, double *, uint64_t *);

template <bool LABEL, bool SUMS, bool MASS, bool TILE>
PassKernel pick_pass(int dmax)
{
    switch (dmax) {
    case 1: return f64_pass_kernel<1, LABEL, SUMS, MASS, TILE>;
    case 2: return f64_pass_kernel<2, LABEL, SUMS, MASS, TILE>;
    case 4: return f64_pass_kernel<4, LABEL, SUMS, MASS, TILE>;
    case 8: return f64_pass_kernel<8, LABEL, SUMS, MASS, TILE>;
    case 16: return f64_pass_kernel<16, LABEL, SUMS, MASS, TILE>;
    case 32: return f64_pass_kernel<32, LABEL, SUMS, MASS, TILE>;
    default: return nullptr;
    }
}

template <bool LABEL, bool SUMS, bool MASS>
PassKernel pick_pass_tile(int dmax, bool tile)
{
    return tile ? pick_pass<LABEL, SUMS, MASS, true>(dmax) : pick_pass<LABEL, SUMS, MASS, false>(dmax);
}

// Register bucket of the FP64 pass: 1..32, 0 beyond.
int dmax_f64(uint64_t d)
{
    const int b = dmax_bucket(d);
    return b > 32 ? 0 : b;
}

int check_f64_sizes(uint64_t d, uint64_t n, uint64_t k)
{
    CLK_REQUIRE(d >= 1 && k >= 1, CLK_ERR_SIZE, "d and k must be >= 1");
    CLK_REQUIRE(n < (1ull << 40), CLK_ERR_SIZE, "n too large");
    CLK_REQUIRE(k <= 0xFFFFFFFFull && d <= 0xFFFFull, CLK_ERR_SIZE, "k or d too large");
    return CLK_OK;
}

// One pass: labels (LABEL) and/or accumulation into sums / masses.
int run_f64_pass(clk_context *ctx, const clk_kernel_config *cfg, bool label, bool want_sums,
                 bool want_mass, uint64_t d, uint64_t n, uint64_t k, const double *points,
                 const double *centroids, uint64_t *labels, double *sums, uint64_t *masses,
                 hipStream_t stream, const char *name)
{
    if (n == 0)
        return CLK_OK;
    const uint64_t b = cfg && cfg->local_size[0] ? cfg->local_size[0] : 256;
    CLK_REQUIRE(b >= 1 && b <= 1024 && b <= (uint64_t)ctx->max_threads_per_block, CLK_ERR_INVALID_ARGUMENT,
                "local_size must be 1..1024");
    const int B = (int)b;
    const int dmax = dmax_f64(d);
    const uint64_t cen_bytes = label ? k * (uint64_t)std::max(dmax, 1) * 8 : 0;
    const uint64_t tile_bytes = (want_sums ? d * k * 8 : 0) + (want_mass ? k * 4 : 0);
    const bool fits_cen = dmax != 0 && cen_bytes <= ctx->lds_per_block;
    const bool tile = fits_cen && cen_bytes + tile_bytes <= ctx->lds_per_block;
    PassKernel kern = nullptr;
    if (fits_cen) {
        const int sel = (label ? 4 : 0) | (want_sums ? 2 : 0) | (want_mass ? 1 : 0);
        switch (sel) {
        case 7: kern = pick_pass_tile<true, true, true>(dmax, tile); break;
        case 4: kern = pick_pass_tile<true, false, false>(dmax, tile); break;
        case 2: kern = pick_pass_tile<false, true, false>(dmax, tile); break;
        case 1: kern = pick_pass_tile<false, false, true>(dmax, tile); break;
        default: break;
        }
    }
    const size_t lds = tile ? (size_t)(cen_bytes + tile_bytes) : (size_t)cen_bytes;
    const uint64_t need = div_up(n, (uint64_t)B);
    uint64_t grid = (cfg && cfg->global_size[0]) ? div_up(cfg->global_size[0], (uint64_t)B)
                    : kern ? auto_grid(ctx, (const void *)kern, B, lds, n)
                           : std::min<uint64_t>(need, (uint64_t)ctx->num_cus * 8);
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, 1u << 16));
    KernelTimer t(ctx, name, stream);
    if (kern) {
        double *ps = nullptr;
        uint32_t *pm = nullptr;
        if (tile && want_sums) {
            CLK_TRY(ensure_scratch(ctx->part_sums, grid * d * k * sizeof(double)));
            ps = static_cast<double *>(ctx->part_sums.ptr);
        }
        if (tile && want_mass) {
            CLK_TRY(ensure_scratch(ctx->part_masses, grid * k * sizeof(uint32_t)));
            pm = static_cast<uint32_t *>(ctx->part_masses.ptr);
        }
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(B), lds, stream, points, n, (uint32_t)d,
                           (uint32_t)k, centroids, labels, ps, pm, sums, masses);
        CLK_HIP_CHECK(hipGetLastError());
        if (tile && want_sums)
            CLK_TRY(launch_reduce_f64(ctx, ps, grid, d * k, sums, true, stream));
        if (tile && want_mass) {
            hipLaunchKernelGGL(reduce_masses_u64_kernel, dim3((unsigned)div_up(k, 256)), dim3(256), 0, stream,
                               pm, grid, (uint32_t)k, masses);
            CLK_HIP_CHECK(hipGetLastError());
        }
        return CLK_OK;
    }
    // beyond the register buckets or the LDS: global memory and atomics
    const int sel = (label ? 4 : 0) | (want_sums ? 2 : 0) | (want_mass ? 1 : 0);
    const dim3 g((unsigned)grid), bl(B);
    switch (sel) {
    case 7:
        hipLaunchKernelGGL((f64_pass_global_kernel<true, true, true>), g, bl, 0, stream, points, n, (uint32_t)d,
                           (uint32_t)k, centroids, labels, sums, masses);
        break;
    case 4:
        hipLaunchKernelGGL((f64_pass_global_kernel<true, false, false>), g, bl, 0, stream, points, n,
                           (uint32_t)d, (uint32_t)k, centroids, labels, sums, masses);
        break;
    case 2:
        hipLaunchKernelGGL((f64_pass_global_kernel<false, true, false>), g, bl, 0, stream, points, n,
                           (uint32_t)d, (uint32_t)k, centroids, labels, sums, masses);
        break;
    case 1:
        hipLaunchKernelGGL((f64_pass_global_kernel<false, false, true>), g, bl, 0, stream, points, n,
                           (uint32_t)d, (uint32_t)k, centroids, labels, sums, masses);
        break;
    default:
        break;
    }
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

int require_stage(const clk_kernel_config *cfg, int stage, const char *what)
{
    CLK_REQUIRE(!cfg || cfg->stage == stage, CLK_ERR_INVALID_ARGUMENT, what);
    return CLK_OK;
}

}  // namespace
}  // namespace clk

using namespace clk;

extern "C" int clk_labeling_f64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                                uint64_t k, const double *points, const double *centroids, uint64_t *labels,
                                void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_TRY(require_stage(cfg, CLK_STAGE_LABELING, "clk_labeling_f64: config is not a labeling-stage config"));
    CLK_TRY(check_f64_sizes(d, n, k));
    CLK_REQUIRE(n == 0 || (points && centroids && labels), CLK_ERR_INVALID_ARGUMENT, "labeling_f64: null buffer");
    return run_f64_pass(ctx, cfg, true, false, false, d, n, k, points, centroids, labels, nullptr, nullptr,
                        as_stream(stream), "LabelingUnrollVector");
}

extern "C" int clk_mass_update_u64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t n, uint64_t k,
                                   const uint64_t *labels, uint64_t *masses, void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_TRY(require_stage(cfg, CLK_STAGE_MASS_UPDATE, "clk_mass_update_u64: config is not a mass-update config"));
    CLK_TRY(check_f64_sizes(1, n, k));
    CLK_REQUIRE(n == 0 || (labels && masses), CLK_ERR_INVALID_ARGUMENT, "mass_update_u64: null buffer");
    return run_f64_pass(ctx, cfg, false, false, true, 1, n, k, nullptr, nullptr, const_cast<uint64_t *>(labels),
                        nullptr, masses, as_stream(stream), "MassUpdate");
}

extern "C" int clk_centroid_update_f64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                                       uint64_t k, const double *points, double *sums, const uint64_t *labels,
                                       const uint64_t *masses, void *stream)
{
    (void)masses;  // accepted and unused, as in the reference kernels
    CLK_TRY(enter(ctx));
    CLK_TRY(require_stage(cfg, CLK_STAGE_CENTROID_UPDATE,
                          "clk_centroid_update_f64: config is not a centroid-update config"));
    CLK_TRY(check_f64_sizes(d, n, k));
    CLK_REQUIRE(n == 0 || (points && sums && labels), CLK_ERR_INVALID_ARGUMENT, "centroid_update_f64: null buffer");
    return run_f64_pass(ctx, cfg, false, true, false, d, n, k, points, nullptr, const_cast<uint64_t *>(labels),
                        sums, nullptr, as_stream(stream), "CentroidUpdate");
}

extern "C" int clk_fused_f64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n, uint64_t k,
                             const double *points, const double *old_centroids, double *sums, uint64_t *labels,
                             uint64_t *masses, void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_TRY(require_stage(cfg, CLK_STAGE_FUSED, "clk_fused_f64: config is not a fused-stage config"));
    CLK_TRY(check_f64_sizes(d, n, k));
    CLK_REQUIRE(cfg == nullptr || cfg->strategy == CLK_FUSED_CLUSTER_MERGE || cfg->strategy == CLK_FUSED_FEATURE_SUM ||
                    cfg->strategy == CLK_FUSED_CLUSTER_MERGE_BOUNDED,
                CLK_ERR_INVALID_ARGUMENT, "fused_f64: unknown strategy id");
    CLK_REQUIRE(n == 0 || (points && old_centroids && sums && labels && masses), CLK_ERR_INVALID_ARGUMENT,
                "fused_f64: null buffer");
    return run_f64_pass(ctx, cfg, true, true, true, d, n, k, points, old_centroids, labels, sums, masses,
                        as_stream(stream), "FusedClusterMerge");
}

extern "C" int clk_matrix_divide_f64(clk_context *ctx, uint64_t d, uint64_t k, const double *sums,
                                     const uint64_t *masses, double *centroids, const double *old_centroids,
                                     double *max_shift, void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_TRY(check_f64_sizes(d, 0, k));
    CLK_REQUIRE(sums && masses && centroids, CLK_ERR_INVALID_ARGUMENT, "matrix_divide_f64: null buffer");
    hipStream_t s = as_stream(stream);
    KernelTimer t(ctx, "FinalizeDivide", s);
    hipLaunchKernelGGL(divide_f64_kernel, dim3(1), dim3(256), 0, s, sums, masses, (uint32_t)d, (uint32_t)k,
                       centroids, max_shift ? old_centroids : nullptr, max_shift);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

extern "C" int clk_lloyd_step_f64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d, uint64_t n,
                                  uint64_t k, const double *points, const double *old_centroids,
                                  double *new_centroids, double *sums, uint64_t *labels, uint64_t *masses,
                                  double *max_shift, void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(new_centroids && new_centroids != old_centroids, CLK_ERR_INVALID_ARGUMENT,
                "lloyd_step_f64: new_centroids must be set and must not alias old_centroids");
    CLK_REQUIRE(sums && masses, CLK_ERR_INVALID_ARGUMENT, "lloyd_step_f64: null buffer");
    CLK_TRY(clk_fill_zero(ctx, sums, d * k * sizeof(double), stream));
    CLK_TRY(clk_fill_zero(ctx, masses, k * sizeof(uint64_t), stream));
    CLK_TRY(clk_fused_f64(ctx, cfg, d, n, k, points, old_centroids, sums, labels, masses, stream));
    return clk_matrix_divide_f64(ctx, d, k, sums, masses, new_centroids, old_centroids, max_shift, stream);
}

extern "C" int clk_matrix_add_u64(clk_context *ctx, uint64_t count, uint64_t *fst, const uint64_t *snd,
                                  void *stream)
{
    CLK_TRY(enter(ctx));
    if (count == 0)
        return CLK_OK;
    CLK_REQUIRE(fst && snd, CLK_ERR_INVALID_ARGUMENT, "matrix_add_u64: null buffer");
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(add_u64_kernel, dim3((unsigned)std::min<uint64_t>(div_up(count, 256), 4096)), dim3(256), 0,
                       s, fst, snd, count);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}
