// Exact centroid sums.
//
// The reference adds FP32 coordinates in an order fixed by its kernel
// geometry (lloyd_fused_cluster_merge.cl's per-work-item tiles, then
// reduce_vector_parcol's tree), so its sums -- and the centroids, and after a
// near tie the labels -- depend on that geometry.  This library makes the
// sums independent of every execution detail instead: each coordinate is
// first rounded (to nearest, ties to even) to an integer multiple of a
// per-feature power of two, the unit, and the integers are added exactly.
// Integer addition is associative, so the fused pass, the three-stage
// pipeline, the chunked (_buffered) pipeline, any grid or tile size and the
// batch engine all produce the same bits for the same (points, labels).
//
// The grid of feature f is a function of the data only: with
//   M_f = max |x_f| over the finite coordinates of the job's points,
//   E_f = the least integer with M_f < 2^E_f (at least -100; -100 when M_f
//         is 0 or subnormal),
//   b   = min(25, 53 - ceil(log2(n + 1))) for the job's n points,
// the unit is 2^(E_f - b) and q(x) = rint(x * 2^(b - E_f)) satisfies
// |q| < 2^b <= 2^25.  So:
//   - 64 values of one cell fit an int32 (the private LDS columns flush to
//     a shared int64 accumulator at least every 64 values),
//   - any sum of at most n values has |sum| < 2^53 units, so it is exact as
//     an int64 and as an FP64 multiple of the unit: the FP64 partial rows and
//     their merges in any order are exact.
// The rounding moves a coordinate by at most 2^(E_f - b - 1) <= 2^-26 M_f
// (about 1.5e-8 of the feature's largest magnitude).
//
// Values outside the grid -- NaN, +-inf, or (with a grid pinned on the
// context for other data) |q| >= 2^25 -- are not added; a per-cell flag word
// records them and the cell's sum becomes NaN, +inf or -inf, as an FP32 or
// FP64 sum containing them would in any order (NaN if a NaN or both
// infinities occur, else the infinity; an out-of-grid finite value is
// reported as NaN).
#pragma once

#include "clk_common.hpp"

namespace clk {

// |x * scale| must stay below this (2^25) for the value to be on the grid.
constexpr float kGridLimit = 33554432.0f;
constexpr int kGridMaxBits = 25;
constexpr int kGridMinExp = -100;

// Kernel argument: per-feature max |x| bit patterns (finite values; the bits
// of a non-negative float order like its value) and b.  gmax[d] is non-zero
// when the grid's data hold a NaN or an infinity (or when it is not known).
// own: the grid was computed from the very points the kernel reads (per
// call, or pinned from the same buffer and n); then, when gmax[d] is 0,
// every coordinate is finite and on the grid and needs no range check.
struct GridRef {
    const uint32_t *gmax;
    uint32_t bits;
    uint32_t own = 0;
};

// b = min(25, 53 - ceil(log2(n + 1))).
__host__ __device__ inline uint32_t grid_bits(uint64_t n)
{
    uint32_t lg = 0;  // ceil(log2(n + 1)) = bit length of n
    while (lg < 64 && (n >> lg) != 0)
        ++lg;
    const int b = 53 - (int)lg;
    return (uint32_t)(b < kGridMaxBits ? b : kGridMaxBits);
}

// E_f: the least integer with M < 2^E, at least -100, from the bits of M.
__host__ __device__ inline int grid_exp(uint32_t maxbits)
{
    const int ex = (int)((maxbits >> 23) & 0xFFu);
    const int e = ex == 0 ? kGridMinExp : ex - 126;
    return e < kGridMinExp ? kGridMinExp : e;
}

// 2^(b - E) as a float (b - E lies in [-104, 125]: a normal float).
__host__ __device__ inline float grid_scale(uint32_t maxbits, uint32_t bits)
{
    const int s = (int)bits - grid_exp(maxbits);
    const uint32_t u = (uint32_t)(s + 127) << 23;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// The unit 2^(E - b) as a double (1 / scale, exact).
__host__ __device__ inline double grid_unit(float scale) { return 1.0 / (double)scale; }

// The grid integer of x and whether x is on the grid.  y = x * scale is
// exact (scale is a power of two and |y| < 2^25 when on the grid); rint
// rounds to nearest, ties to even.  Off the grid (NaN, inf, too large) the
// result is 0.
__device__ __forceinline__ int32_t grid_q(float x, float scale, bool &ok)
{
    const float y = x * scale;
    ok = __builtin_fabsf(y) < kGridLimit;
    const int32_t q = (int32_t)__builtin_rintf(y);
    return ok ? q : 0;
}

// The grid integer of a value known to be on the grid (data whose grid was
// computed from themselves and hold no NaN / inf).
__device__ __forceinline__ int32_t grid_q_on(float x, float scale)
{
    return (int32_t)__builtin_rintf(x * scale);
}

// Flag bits of a value off the grid: 1 NaN (or finite but off the grid),
// 2 +inf, 4 -inf.
__host__ __device__ inline uint32_t special_flag(float x)
{
    return x == __builtin_inff() ? 2u : x == -__builtin_inff() ? 4u : 1u;
}

// The sum of a cell whose flags are non-zero.
__host__ __device__ inline double special_value(uint32_t flags)
{
    if ((flags & 1u) || (flags & 6u) == 6u)
        return __builtin_nan("");
    return (flags & 2u) ? __builtin_inf() : -__builtin_inf();
}

// A cell's FP64 sum: the exact integer sum times the unit, or the special
// value.  (double)q is exact: |q| < 2^53.
__host__ __device__ inline double cell_value(int64_t q, double unit, uint32_t flags)
{
    return flags ? special_value(flags) : (double)q * unit;
}

// LDS int64 / uint32 atomics without return value (ds_add_u64, ds_or_b32):
// exact, so the order in which work-items add does not matter.
__device__ __forceinline__ void lds_add64(long long *p, long long v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_or(uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Host side: the grid of (points [d][n]) into gmax[d + 1] (device), on
// stream: zero, then the max-|x| pass (gmax[d]: NaN / inf seen).
int launch_grid_max(clk_context *ctx, uint64_t d, uint64_t n, const float *points, uint32_t *gmax,
                    hipStream_t stream);

// The grid the sum-accumulating ops of ctx use for (d, n, points): the
// pinned one (clk_sum_grid_from_*) when set, else computed now into the
// context's scratch.  out->own: computed now, or pinned from these points.
int resolve_grid(clk_context *ctx, uint64_t d, uint64_t n, const float *points, hipStream_t stream,
                 GridRef *out);

}  // namespace clk
