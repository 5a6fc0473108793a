// Mass update: masses[c] += #{p : labels[p] == c}, four strategies with the
// reference's names (histogram_global.cl, histogram_part_global.cl,
// histogram_part_local.cl, histogram_part_private.cl), built for gfx950:
//   global_atomic  wave-aggregated atomics straight into `masses`
//   part_global    per-work-group rows of a global scratch histogram
//                  (global atomics), then a fixed-order column reduction
//   part_local     one LDS histogram per work-group (ds_add_u32), partials
//   part_private   one LDS histogram column per work-item (no collisions),
//                  summed per work-group, partials
// Integer arithmetic: every strategy gives identical, exact masses.  Labels
// >= k are skipped.
#include "clk_common.hpp"
#include "clk_kernels.hpp"

#include <algorithm>

namespace clk {
namespace {

template <int VEC>
__global__ void __launch_bounds__(1024)
mass_global_atomic_kernel(const uint32_t *__restrict__ labels, uint64_t n, uint32_t k,
                          uint32_t *__restrict__ masses)
{
    const uint64_t groups = div_up(n, VEC);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x % kWave;
    for (uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x; g0 < groups; g0 += stride) {
        const uint64_t g = g0 + threadIdx.x;
        uint32_t lab[VEC];
        bool ok[VEC];
        const uint64_t p = g * VEC;
        if (g < groups && p + VEC <= n) {
            load_labels<VEC>(labels, p, lab);
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                ok[v] = lab[v] < k;
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                ok[v] = (g < groups) && (p + v < n);
                lab[v] = ok[v] ? labels[p + v] : 0u;
                ok[v] = ok[v] && lab[v] < k;
            }
        }
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            // One atomic per distinct label present in the wave.
            bool active = ok[v];
            while (__any(active)) {
                const uint64_t act = __ballot(active);
                const int leader = __ffsll((unsigned long long)act) - 1;
                const uint32_t l = __shfl(lab[v], leader);
                const uint64_t same = __ballot(active && lab[v] == l);
                if (lane == (uint32_t)leader)
                    atomicAdd(&masses[l], (uint32_t)__popcll(same));
                if (active && lab[v] == l)
                    active = false;
            }
        }
    }
}

// Work-group row of a global histogram, updated with global atomics.
__global__ void __launch_bounds__(1024)
mass_part_global_kernel(const uint32_t *__restrict__ labels, uint64_t n, uint32_t k,
                        uint32_t *__restrict__ part)
{
    uint32_t *row = part + (uint64_t)blockIdx.x * k;
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x)
        row[c] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const uint32_t c = labels[p];
        if (c < k)
            atomicAdd(&row[c], 1u);
    }
}

// PRIVATE: LDS [k][blockDim] (column per work-item); else LDS [k] shared.
template <int VEC, bool PRIVATE>
__global__ void __launch_bounds__(1024)
mass_lds_kernel(const uint32_t *__restrict__ labels, uint64_t n, uint32_t k,
                uint32_t *__restrict__ part)
{
    extern __shared__ __align__(16) uint32_t hist[];
    const uint32_t B = blockDim.x, tid = threadIdx.x;
    const uint32_t cols = PRIVATE ? B : 1;
    for (uint32_t i = tid; i < k * cols; i += B)
        hist[i] = 0;
    __syncthreads();
    uint32_t *mine = hist + (PRIVATE ? tid : 0);
    const uint64_t groups = n / VEC;
    const uint64_t stride = (uint64_t)gridDim.x * B;
    for (uint64_t g = (uint64_t)blockIdx.x * B + tid; g < groups; g += stride) {
        uint32_t lab[VEC];
        load_labels<VEC>(labels, g * VEC, lab);
#pragma unroll
        for (int v = 0; v < VEC; ++v)
            if (lab[v] < k) {
                if (PRIVATE)
                    lds_acc(mine + __umul24(lab[v], cols), 1u);
                else
                    lds_add(mine + __umul24(lab[v], cols), 1u);
            }
    }
    // Tail (n % VEC points) by work-group 0.
    if (blockIdx.x == 0) {
        for (uint64_t p = groups * VEC + tid; p < n; p += B) {
            const uint32_t c = labels[p];
            if (c < k) {
                if (PRIVATE)
                    lds_acc(mine + __umul24(c, cols), 1u);
                else
                    lds_add(mine + __umul24(c, cols), 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t c = tid; c < k; c += B)
        part[(uint64_t)blockIdx.x * k + c] =
            PRIVATE ? lds_row_sum_u32(hist + (uint64_t)c * cols, cols, c) : hist[c];
}

using MassKernel = void (*)(const uint32_t *, uint64_t, uint32_t, uint32_t *);

MassKernel pick_lds(int vec, bool priv)
{
    if (priv)
        return vec == 4 ? mass_lds_kernel<4, true> : vec == 2 ? mass_lds_kernel<2, true>
                                                               : mass_lds_kernel<1, true>;
    return vec == 4 ? mass_lds_kernel<4, false> : vec == 2 ? mass_lds_kernel<2, false>
                                                            : mass_lds_kernel<1, false>;
}

}  // namespace

int run_mass_update(clk_context *ctx, const clk_kernel_config *cfg, uint64_t n, uint64_t k,
                    const uint32_t *labels, uint32_t *masses, hipStream_t stream)
{
    CLK_REQUIRE(k >= 1 && k <= 0xFFFFFFFFull, CLK_ERR_SIZE, "mass_update: bad k");
    CLK_REQUIRE(labels && masses, CLK_ERR_INVALID_ARGUMENT, "mass_update: null buffer");
    if (n == 0)
        return CLK_OK;
    int strategy = cfg ? cfg->strategy : CLK_MASS_PART_PRIVATE;
    const uint64_t B = (cfg && cfg->local_size[0]) ? cfg->local_size[0] : 256;
    CLK_REQUIRE(B <= (uint64_t)ctx->max_threads_per_block, CLK_ERR_INVALID_ARGUMENT,
                "mass_update: local_size exceeds the device limit");
    const int vec = pick_vector_width(cfg ? cfg->vector_length : 4);
    const uint64_t user_grid = (cfg && cfg->global_size[0]) ? div_up(cfg->global_size[0], B) : 0;

    // LDS-fit fallbacks: private -> local -> part_global.
    if (strategy == CLK_MASS_PART_PRIVATE && k * B * 4 > ctx->lds_per_block)
        strategy = CLK_MASS_PART_LOCAL;
    if (strategy == CLK_MASS_PART_LOCAL && k * 4 > ctx->lds_per_block)
        strategy = CLK_MASS_PART_GLOBAL;

    switch (strategy) {
    case CLK_MASS_GLOBAL_ATOMIC: {
        MassKernel kern = vec == 4 ? mass_global_atomic_kernel<4>
                        : vec == 2 ? mass_global_atomic_kernel<2>
                                   : mass_global_atomic_kernel<1>;
        uint64_t grid = user_grid ? user_grid
                                  : auto_grid(ctx, (const void *)kern, (int)B, 0, div_up(n, vec));
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, 1u << 20));
        KernelTimer t(ctx, "MassUpdateGlobalAtomic", stream);
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3((unsigned)B), 0, stream, labels, n,
                           (uint32_t)k, masses);
        CLK_HIP_CHECK(hipGetLastError());
        return CLK_OK;
    }
    case CLK_MASS_PART_GLOBAL:
    case CLK_MASS_PART_LOCAL:
    case CLK_MASS_PART_PRIVATE: {
        MassKernel kern;
        size_t lds = 0;
        const char *name;
        if (strategy == CLK_MASS_PART_GLOBAL) {
            kern = mass_part_global_kernel;
            name = "MassUpdatePartGlobal";
        } else if (strategy == CLK_MASS_PART_LOCAL) {
            kern = pick_lds(vec, false);
            lds = k * 4;
            name = "MassUpdatePartLocal";
        } else {
            kern = pick_lds(vec, true);
            lds = k * B * 4;
            name = "MassUpdatePartPrivate";
        }
        uint64_t grid = user_grid ? user_grid
                                  : auto_grid(ctx, (const void *)kern, (int)B, lds, div_up(n, vec));
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, 1u << 20));
        CLK_TRY(ensure_scratch(ctx->part_masses, grid * k * sizeof(uint32_t)));
        uint32_t *pm = static_cast<uint32_t *>(ctx->part_masses.ptr);
        {
            KernelTimer t(ctx, name, stream);
            hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3((unsigned)B), lds, stream, labels,
                               n, (uint32_t)k, pm);
            CLK_HIP_CHECK(hipGetLastError());
        }
        return launch_reduce_u32(ctx, pm, grid, k, masses, true, stream);
    }
    default:
        set_error("mass_update: unknown strategy id");
        return CLK_ERR_INVALID_ARGUMENT;
    }
}

}  // namespace clk

extern "C" int clk_mass_update(clk_context *ctx, const clk_kernel_config *cfg, uint64_t n,
                               uint64_t k, const uint32_t *labels, uint32_t *masses,
                               void *stream)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(!cfg || cfg->stage == CLK_STAGE_MASS_UPDATE, CLK_ERR_INVALID_ARGUMENT,
                "clk_mass_update: config is not a mass-update-stage config");
    return clk::run_mass_update(ctx, cfg, n, k, labels, masses, clk::as_stream(stream));
}
