// The sum grid (clk_grid.hpp): per-feature max |x| pass, the context's
// pinned grid and its C ABI.
#include "clk_grid.hpp"
#include "clk_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace clk {
namespace {

// blockIdx.y = feature; the work-items of blockIdx.x stride over the row.
// A wave's maximum goes to gmax[f] with one atomicMax (bits of non-negative
// floats order like their values).  NaN and +-inf are skipped.
__global__ void __launch_bounds__(256)
grid_max_kernel(const float *__restrict__ pts, uint64_t n, uint32_t *__restrict__ gmax, uint32_t d)
{
    const uint32_t f = blockIdx.y;
    const float *row = pts + (uint64_t)f * n;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t m = 0;
    bool special = false;  // a NaN or an infinity seen
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto take = [&](float x) {
        const float a = __builtin_fabsf(x);
        const uint32_t u = __float_as_uint(a);
        const bool fin = a <= FLT_MAX;
        m = (fin && u > m) ? u : m;
        special = special || !fin;
    };
    // 128-bit loads (4-byte aligned rows), then the n % 4 tail values
    const gf32x4 *r4 = reinterpret_cast<const gf32x4 *>(row);
    for (uint64_t i = t0; i < n / 4; i += stride) {
        const gf32x4 v = stream_load(r4 + i);
        take(v.x);
        take(v.y);
        take(v.z);
        take(v.w);
    }
    if (t0 < n % 4)
        take(row[n / 4 * 4 + t0]);
    for (int o = 32; o >= 1; o /= 2) {
        const uint32_t other = (uint32_t)__shfl_xor((int)m, o);
        m = other > m ? other : m;
    }
    if (threadIdx.x % 64 == 0 && m != 0)
        atomicMax(&gmax[f], m);
    if (__ballot(special) != 0 && threadIdx.x % 64 == 0)
        atomicOr(&gmax[d], 1u);
}

}  // namespace

int launch_grid_max(clk_context *ctx, uint64_t d, uint64_t n, const float *points, uint32_t *gmax,
                    hipStream_t stream)
{
    CLK_TRY(clk_fill_zero(ctx, gmax, (d + 1) * sizeof(uint32_t), stream));
    if (n == 0)
        return CLK_OK;
    CLK_REQUIRE(d <= 65535, CLK_ERR_SIZE, "sum grid: d too large");
    const uint64_t per_row = std::max<uint64_t>(1, (4 * (uint64_t)ctx->num_cus + d - 1) / d);
    const uint64_t gx = std::min<uint64_t>(per_row, div_up(n, 256 * 4));
    KernelTimer t(ctx, "SumGrid", stream);
    hipLaunchKernelGGL(grid_max_kernel, dim3((unsigned)std::max<uint64_t>(gx, 1), (unsigned)d), dim3(256),
                       0, stream, points, n, gmax, (uint32_t)d);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

int resolve_grid(clk_context *ctx, uint64_t d, uint64_t n, const float *points, hipStream_t stream,
                 GridRef *out)
{
    if (ctx->grid_pinned) {
        CLK_REQUIRE(ctx->grid_d == d, CLK_ERR_INVALID_ARGUMENT,
                    "the sum grid pinned on the context has another number of features");
        out->gmax = static_cast<const uint32_t *>(ctx->grid.ptr);
        out->bits = grid_bits(ctx->grid_n);
        out->own = ctx->grid_src != nullptr && ctx->grid_src == points && ctx->grid_n == n;
        return CLK_OK;
    }
    CLK_TRY(ensure_scratch(ctx->grid, (d + 1) * sizeof(uint32_t)));
    CLK_TRY(launch_grid_max(ctx, d, n, points, static_cast<uint32_t *>(ctx->grid.ptr), stream));
    out->gmax = static_cast<const uint32_t *>(ctx->grid.ptr);
    out->bits = grid_bits(n);
    out->own = 1;
    return CLK_OK;
}

}  // namespace clk

using namespace clk;

extern "C" int clk_sum_grid_from_points(clk_context *ctx, uint64_t d, uint64_t n, const float *points,
                                        void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(d >= 1 && d <= 65535, CLK_ERR_SIZE, "clk_sum_grid_from_points: bad d");
    CLK_REQUIRE(points || n == 0, CLK_ERR_INVALID_ARGUMENT, "clk_sum_grid_from_points: null points");
    ctx->grid_pinned = false;
    CLK_TRY(ensure_scratch(ctx->grid, (d + 1) * sizeof(uint32_t)));
    CLK_TRY(launch_grid_max(ctx, d, n, points, static_cast<uint32_t *>(ctx->grid.ptr), as_stream(stream)));
    ctx->grid_pinned = true;
    ctx->grid_d = d;
    ctx->grid_n = n;
    ctx->grid_src = points;
    return CLK_OK;
}

extern "C" int clk_sum_grid_from_max(clk_context *ctx, uint64_t d, uint64_t n, const float *host_max_abs)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(d >= 1 && d <= 65535, CLK_ERR_SIZE, "clk_sum_grid_from_max: bad d");
    CLK_REQUIRE(host_max_abs != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_sum_grid_from_max: null maxima");
    std::vector<uint32_t> bits(d + 1, 0u);
    for (uint64_t f = 0; f < d; ++f) {
        const float a = std::fabs(host_max_abs[f]);
        uint32_t u = 0;
        if (a <= FLT_MAX)
            std::memcpy(&u, &a, 4);
        bits[f] = u;
    }
    bits[d] = 1u;  // whether the data hold NaN / inf is not known: check
    ctx->grid_pinned = false;
    CLK_TRY(ensure_scratch(ctx->grid, (d + 1) * sizeof(uint32_t)));
    CLK_HIP_CHECK(hipMemcpy(ctx->grid.ptr, bits.data(), (d + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    ctx->grid_pinned = true;
    ctx->grid_d = d;
    ctx->grid_n = n;
    ctx->grid_src = nullptr;
    return CLK_OK;
}

extern "C" int clk_sum_grid_clear(clk_context *ctx)
{
    CLK_REQUIRE(ctx != nullptr, CLK_ERR_INVALID_ARGUMENT, "null clk_context");
    ctx->grid_pinned = false;
    ctx->grid_src = nullptr;
    return CLK_OK;
}

extern "C" int clk_sum_grid_units(clk_context *ctx, uint64_t d, uint64_t n, const float *points,
                                  double *host_units, void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(host_units != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_sum_grid_units: null output");
    CLK_REQUIRE(d >= 1 && d <= 65535, CLK_ERR_SIZE, "clk_sum_grid_units: bad d");
    hipStream_t s = as_stream(stream);
    GridRef g;
    CLK_TRY(resolve_grid(ctx, d, n, points, s, &g));
    std::vector<uint32_t> bits(d);
    CLK_HIP_CHECK(hipMemcpyAsync(bits.data(), g.gmax, d * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    CLK_HIP_CHECK(hipStreamSynchronize(s));
    for (uint64_t f = 0; f < d; ++f)
        host_units[f] = grid_unit(grid_scale(bits[f], g.bits));
    return CLK_OK;
}
