// Device building blocks of the MFMA (16-bit operand) distance kernels:
// operand layouts, score tiles, best/second key ranking, the per-point error
// bound and the point loader.  The derivation of the bound and of the margin
// test is in the header comment of clk_mfma.hip; the batch engine
// (clk_batch.hip) uses the same pieces per job.
#pragma once

#include "clk_common.hpp"

#include <cmath>

namespace clk {
namespace mfma {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32v4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int kTile = 16;    // centroids per MFMA row tile
constexpr int kKBlock = 32;  // features per MFMA K block
constexpr int kGroups = 4;   // 16-point column groups per wave (64 points)
constexpr int kMaxKb = 4;    // d <= 128
constexpr float kU = 5.9604645e-8f;  // 2^-24

struct MfmaParams {
    float mu;      // absolute input error bound of a flushed subnormal
    float gamma;   // MFMA summation bound factor
    float trunc;   // relative truncation of the key, 2^(KB-23)
    float eps;     // exact-distance relative error bound
    float sqrt_d;  // sqrt(d), rounded up
    uint32_t key_bits;
};

template <bool FP16>
__device__ __forceinline__ uint16_t to16(float x)
{
    if constexpr (FP16)
        return __builtin_bit_cast(uint16_t, (_Float16)x);
    else
        return __builtin_bit_cast(uint16_t, (__bf16)x);
}

template <bool FP16>
__device__ __forceinline__ float from16(uint16_t v)
{
    if constexpr (FP16)
        return (float)__builtin_bit_cast(_Float16, v);
    else
        return __uint_as_float((uint32_t)v << 16);
}

template <bool FP16>
__device__ __forceinline__ f32v4 mma16(u16x8 a, u16x8 b, f32v4 c)
{
    if constexpr (FP16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Value of lane (lane ^ O) for O < 32, by ds_swizzle (bit mode: and 0x1F,
// or 0, xor O).  Unlike __shfl_xor it needs no per-lane address register.
template <int O>
__device__ __forceinline__ uint32_t xor_swizzle(uint32_t v)
{
    static_assert(O > 0 && O < 32, "ds_swizzle bit mode stays within 32 lanes");
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (O << 10));
}

template <int O>
__device__ __forceinline__ float xor_swizzle(float v)
{
    return __uint_as_float(xor_swizzle<O>(__float_as_uint(v)));
}

__device__ __forceinline__ uint32_t min3_u32(uint32_t a, uint32_t b, uint32_t c)
{
    return min(a, min(b, c));  // v_min3_u32
}

// Middle value of three (v_med3_u32).  The operands are VALU results, never
// MFMA results, so the asm needs no MFMA-read wait states.
__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Two new keys a, b into (best, second), best <= second: the two smallest of
// {best, second, a, b} are min3(best, a, b) and min(second, med3(best, a, b)).
__device__ __forceinline__ void push2(uint32_t &best, uint32_t &second, uint32_t a, uint32_t b)
{
    const uint32_t m = med3_u32(best, a, b);
    best = min3_u32(best, a, b);
    second = min(second, m);
}

// Hide known bits of a value from the optimizer (emits no instruction), so
// (score & mask) | index stays one v_and_or_b32: opaque_v pins a value to a
// vector register, opaque_s a wave-uniform value to a scalar register (one
// scalar operand per VALU instruction, so the mask goes in a VGPR and the
// index in an SGPR).
__device__ __forceinline__ uint32_t opaque_v(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ uint32_t opaque_s(uint32_t v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// A operands (one per 32-feature block) and norms of centroid tile t, read
// from LDS ahead of use.
template <int NKB>
struct TileOps {
    u16x8 a[NKB];
    f32v4 nrm;
};

template <int NKB>
__device__ __forceinline__ TileOps<NKB> load_ops(const u16x8 *lds_table, const float *lds_norm,
                                                 uint32_t t, uint32_t lane, uint32_t h)
{
    TileOps<NKB> o;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
        o.a[kb] = lds_table[((uint64_t)t * NKB + kb) * 64 + lane];
    o.nrm = *reinterpret_cast<const f32v4 *>(lds_norm + t * kTile + 4 * h);
    return o;
}

// The wave's 64 points as 16-bit MFMA B operands: lane (q, h) holds, for
// group g, features kb*32 + 8h + j of that group's point.
template <int NKB>
struct PointOps {
    u16x8 b[kGroups][NKB];
};

// Scores of one centroid tile for the wave's 4 point groups: lane (q, h)
// gets rows 4h..4h+3 (centroids t*16 + 4h + i) of point column q.
// ADD_K = false: the norms already hold RN(||c||^2 + K).
template <bool FP16, int NKB, bool ADD_K = true>
__device__ __forceinline__ void score_tile(f32v4 (&acc)[kGroups], const TileOps<NKB> &o, float K,
                                           const PointOps<NKB> &x)
{
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        if constexpr (ADD_K)
            acc[g] = o.nrm + K;
        else
            acc[g] = o.nrm;
    }
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
        for (int g = 0; g < kGroups; ++g)
            acc[g] = mma16<FP16>(o.a[kb], x.b[g][kb], acc[g]);
}

// Keys of one scored tile t into the running (best, second) of each group.
// Inside the ranking a key carries the lane-local index 4t + i of its row
// (i < 4; the lane's rows are centroids 16t + 4h + i): it is the same for
// every lane, so it sits in a scalar register and costs no VALU op, and for a
// fixed h it orders like the centroid index.  rank_all converts the final
// keys to centroid indices.
__device__ __forceinline__ void rank_tile(const f32v4 (&acc)[kGroups], uint32_t t, uint32_t vkeep,
                                          uint32_t (&best)[kGroups], uint32_t (&second)[kGroups])
{
    uint32_t idx[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        idx[i] = opaque_s(4 * t + i);
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        uint32_t key[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            key[i] = (__float_as_uint(acc[g][i]) & vkeep) | idx[i];
        push2(best[g], second[g], key[0], key[1]);
        push2(best[g], second[g], key[2], key[3]);
    }
}

// Lane-local key index 4t + i -> centroid index 16t + 4h + i (fits the key's
// index bits: 16 * ntiles <= 2^key_bits).
__device__ __forceinline__ uint32_t key_to_centroid(uint32_t key, uint32_t kmask, uint32_t h)
{
    const uint32_t l = key & kmask;
    return (key & ~kmask) | ((l & ~3u) << 2) | (h << 2) | (l & 3u);
}

// Merge two (best, second) key pairs into the two smallest of the four.
__device__ __forceinline__ void merge_pair(uint32_t &b, uint32_t &s, uint32_t b2, uint32_t s2)
{
    const uint32_t nb = min(b, b2);
    s = min(max(b, b2), min(s, s2));
    b = nb;
}

// Exact FP32 distance, the reference's fma sequence from 0
// (lloyd_fused_feature_sum.cl:210-239, kmeans_naive.cpp:49-63), from the
// point's FP32 row staged in LDS and the centroid's row of the transposed
// copy (both padded to a multiple of 4 features, read as float4).
__device__ __forceinline__ float exact_dist(const float *xrow, const float *__restrict__ crow,
                                            uint32_t d)
{
    float dist = 0.0f;
    for (uint32_t f0 = 0; f0 < d; f0 += 16) {
        f32v4 xv[4], cv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            xv[j] = *reinterpret_cast<const f32v4 *>(xrow + f0 + 4 * j);
            cv[j] = *reinterpret_cast<const f32v4 *>(crow + f0 + 4 * j);
        }
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) {
            if (f0 + j < d) {
                const float t = xv[j / 4][j % 4] - cv[j / 4][j % 4];
                dist = __builtin_fmaf(t, t, dist);
            }
        }
    }
    return dist;
}

// Load the wave's 64 points and build their operands: lane (q, h) handles,
// for group g, point pid[g] (>= n: none; zeros), features kb*32 + 8h + j
// (0 past d).  Every load address is in bounds (a missing point reads point
// n-1, a missing feature reads feature 0; the value is then replaced by 0).
// X2[g] >= ||x||^2 (an FP32 sum over the 4 lanes of a point: at most P/4 + 2
// roundings, plus underflow).
template <bool FP16, int NKB, bool FULL, bool KEEP>
__device__ __forceinline__ void load_points_core(PointOps<NKB> &x, float (&X2)[kGroups],
                                                 const float *__restrict__ pts, uint64_t n,
                                                 uint32_t d, const uint64_t (&pid)[kGroups],
                                                 uint32_t q, uint32_t h, float *xs,
                                                 float (&keep)[kGroups][NKB * 8])
{
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        const bool valid = pid[g] < n;
        const uint64_t pc = valid ? pid[g] : n - 1;
        const float *row = pts + pc + (uint64_t)(8 * h) * n;  // feature 8h
        float raw[NKB * 8];
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint64_t uoff = (uint64_t)(kb * kKBlock + j) * n;
                if constexpr (FULL) {
                    const float v = row[uoff];
                    raw[kb * 8 + j] = valid ? v : 0.0f;
                } else {
                    const bool fv = (uint32_t)(kb * kKBlock + 8 * h + j) < d;
                    const float v = *(fv ? row + uoff : pts + pc);
                    raw[kb * 8 + j] = (valid && fv) ? v : 0.0f;
                }
            }
        if (xs) {  // row slot 16g + q, features kb*32 + 8h .. +7
            float *row = xs + (uint32_t)(g * 16 + q) * (NKB * kKBlock) + 8 * h;
#pragma unroll
            for (int kb = 0; kb < NKB; ++kb) {
                *reinterpret_cast<f32v4 *>(row + kb * kKBlock) =
                    f32v4{raw[kb * 8 + 0], raw[kb * 8 + 1], raw[kb * 8 + 2], raw[kb * 8 + 3]};
                *reinterpret_cast<f32v4 *>(row + kb * kKBlock + 4) =
                    f32v4{raw[kb * 8 + 4], raw[kb * 8 + 5], raw[kb * 8 + 6], raw[kb * 8 + 7]};
            }
        }
        if constexpr (KEEP) {
#pragma unroll
            for (int j = 0; j < NKB * 8; ++j)
                keep[g][j] = raw[j];
        }
        float s = 0.0f;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            u16x8 vb;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float v = raw[kb * 8 + j];
                s = __builtin_fmaf(v, v, s);
                vb[j] = to16<FP16>(v);
            }
            x.b[g][kb] = vb;
        }
        s += xor_swizzle<16>(s);
        s += __shfl_xor(s, 32);
        X2[g] = s * (1.0f + (float)(2 * NKB * kKBlock + 8) * kU) + 1e-37f;
    }
}

template <bool FP16, int NKB, bool FULL>
__device__ __forceinline__ void load_points(PointOps<NKB> &x, float (&X2)[kGroups],
                                            const float *__restrict__ pts, uint64_t n, uint32_t d,
                                            const uint64_t (&pid)[kGroups], uint32_t q, uint32_t h,
                                            float *xs = nullptr)
{
    float unused[kGroups][NKB * 8];
    load_points_core<FP16, NKB, FULL, false>(x, X2, pts, n, d, pid, q, h, xs, unused);
}

// As load_points, and keep[g][kb*8 + j] = the FP32 value of feature
// kb*32 + 8h + j of group g's point (0 past n or d).
template <bool FP16, int NKB, bool FULL>
__device__ __forceinline__ void load_points_keep(PointOps<NKB> &x, float (&X2)[kGroups],
                                                 const float *__restrict__ pts, uint64_t n,
                                                 uint32_t d, const uint64_t (&pid)[kGroups],
                                                 uint32_t q, uint32_t h,
                                                 float (&keep)[kGroups][NKB * 8])
{
    load_points_core<FP16, NKB, FULL, true>(x, X2, pts, n, d, pid, q, h, nullptr, keep);
}

// The loads of load_points alone, for software pipelining: raw[g][kb*8 + j]
// = feature kb*32 + 8h + j of point pid[g], already 0 for a missing point or
// feature (read from the zero word *zero instead: the address select is paid
// anyway, and build_operands then needs no masks).
// A float in global memory (address space 1).
typedef __attribute__((address_space(1))) const float gfloat1;

template <int NKB>
__device__ __forceinline__ void fetch_points(float (&raw)[kGroups][NKB * 8], const float *__restrict__ pts,
                                             uint64_t n, uint32_t d, const uint64_t (&pid)[kGroups],
                                             uint32_t h, const float *zero)
{
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        const bool valid = pid[g] < n;
        const float *row = pts + (valid ? pid[g] : 0) + (uint64_t)(8 * h) * n;  // feature 8h
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const bool ok = valid && (uint32_t)(kb * kKBlock + 8 * h + j) < d;
                // both candidates are global memory: a global load (a flat
                // load would also count against the LDS wait counter, so
                // every LDS wait of the caller would wait for it)
                raw[g][kb * 8 + j] = *(const gfloat1 *)(ok ? row + (uint64_t)(kb * kKBlock + j) * n : zero);
            }
    }
}

// The operands and X2 of load_points from fetch_points' values.
template <bool FP16, int NKB>
__device__ __forceinline__ void build_operands(const float (&raw)[kGroups][NKB * 8], PointOps<NKB> &x,
                                               float (&X2)[kGroups])
{
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        float s = 0.0f;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            u16x8 vb;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float v = raw[g][kb * 8 + j];
                s = __builtin_fmaf(v, v, s);
                vb[j] = to16<FP16>(v);
            }
            x.b[g][kb] = vb;
        }
        s += xor_swizzle<16>(s);
        s += __shfl_xor(s, 32);
        X2[g] = s * (1.0f + (float)(2 * NKB * kKBlock + 8) * kU) + 1e-37f;
    }
}

// Everything the margin test needs about one wave's points.
struct WaveBounds {
    float K;               // shared score offset
    float beta;            // K-proportional part of the score error
    float E0[kGroups];     // the rest of the score error bound, per point
    float Xn[kGroups];     // >= ||x||
};

// The 16-bit rounding of a point, ||x̃ - x|| <= u16 ||x|| + s16 sqrt(d):
// each coordinate moves by at most u16 |x_f| (round to nearest, u16 = 2^-11
// for FP16, 2^-8 for BF16) or, below the smallest normal, by half the
// subnormal spacing s16 (2^-25 for FP16; BF16's 2^-134, bounded by 1e-38).
// An FP16 coordinate of magnitude >= 65520 becomes infinite: a point whose
// ||x||^2 bound reaches 65520^2 gets an infinite bound (every centroid is a
// candidate), as does every non-finite point.
template <bool FP16>
__device__ __forceinline__ float rounding_bound(float X, float X2, float sqrt_d)
{
    constexpr float u16 = FP16 ? 4.8828125e-4f : 3.90625e-3f;  // 2^-11, 2^-8
    constexpr float s16 = FP16 ? 2.9802322e-8f : 1e-38f;       // 2^-25, > 2^-134
    const float r = (u16 * X + s16 * sqrt_d) * (1.0f + 4.0f * kU);
    if constexpr (FP16)
        return X2 < 4.29287e9f ? r : __builtin_inff();  // 65520^2 = 4.2928704e9
    else
        return r;
}

template <int NKB, bool FP16>
__device__ __forceinline__ WaveBounds wave_bounds(const float (&X2)[kGroups], const float *stats,
                                                  const MfmaParams &prm)
{
    const float Cx = stats[0], Nx = stats[1], R = stats[3];
    const float gm = prm.gamma, ms = prm.mu * prm.sqrt_d;
    const float absP = (float)(NKB * kKBlock + 2) * 2.3509887e-38f;  // (P+2) 2^-125
    WaveBounds w;
    float kneed = 0.0f;
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        const float X = sqrtf(X2[g]) * (1.0f + 4.0f * kU);
        const float Rx = rounding_bound<FP16>(X, X2[g], prm.sqrt_d) + ms;
        const float in = (X + Rx) * (R + ms) + Rx * 2.0f * Cx;
        const float Pm = (X + Rx) * (2.0f * Cx + R + ms);
        w.Xn[g] = X;
        w.E0[g] = 1.01f * (in + gm * (Nx + Pm) + 2.0f * kU * Nx) + absP;
        kneed = fmaxf(kneed, X2[g] + w.E0[g]);  // a NaN point does not poison K
    }
    kneed = fmaxf(kneed, xor_swizzle<1>(kneed));
    kneed = fmaxf(kneed, xor_swizzle<2>(kneed));
    kneed = fmaxf(kneed, xor_swizzle<4>(kneed));
    kneed = fmaxf(kneed, xor_swizzle<8>(kneed));
    kneed = fmaxf(kneed, xor_swizzle<16>(kneed));
    kneed = fmaxf(kneed, __shfl_xor(kneed, 32));
    w.beta = 1.01f * (gm + kU);
    w.K = kneed * (1.0f + 1e-5f) / (1.0f - w.beta) + 1e-30f;
    return w;
}

// Best / second keys over all centroid tiles, one tile at a time, the next
// tile's operands read from LDS while this one is scored and ranked; the
// kernels run several waves per SIMD to cover the MFMA latency.  PIPE: two
// score sets in turn, the next tile's MFMAs issued before this tile's
// ranking, so a wave's own VALU work runs under its matrix work (16 more
// registers: the streaming labeling takes it with three waves per SIMD; the
// batch tiles, already at two per SIMD, do not).
template <bool FP16, int NKB, bool ADD_K = true, bool PIPE = false>
__device__ __forceinline__ void rank_all(const u16x8 *lds_table, const float *lds_norm,
                                         uint32_t ntiles, uint32_t lane, uint32_t h, float K,
                                         uint32_t kmask, const PointOps<NKB> &x,
                                         uint32_t (&best)[kGroups], uint32_t (&second)[kGroups])
{
#pragma unroll
    for (int g = 0; g < kGroups; ++g)
        best[g] = second[g] = 0xFFFFFFFFu;
    const uint32_t vkeep = opaque_v(~kmask);
    if constexpr (PIPE) {
        // two score sets in turn: the next tile's MFMAs are issued before
        // this tile's keys are ranked
        f32v4 A[kGroups], B[kGroups];
        score_tile<FP16, NKB, ADD_K>(A, load_ops<NKB>(lds_table, lds_norm, 0, lane, h), K, x);
        uint32_t t = 0;
        for (; t + 2 < ntiles; t += 2) {
            score_tile<FP16, NKB, ADD_K>(B, load_ops<NKB>(lds_table, lds_norm, t + 1, lane, h), K, x);
            rank_tile(A, t, vkeep, best, second);
            score_tile<FP16, NKB, ADD_K>(A, load_ops<NKB>(lds_table, lds_norm, t + 2, lane, h), K, x);
            rank_tile(B, t + 1, vkeep, best, second);
        }
        if (t + 1 < ntiles) {
            score_tile<FP16, NKB, ADD_K>(B, load_ops<NKB>(lds_table, lds_norm, t + 1, lane, h), K, x);
            rank_tile(A, t, vkeep, best, second);
            rank_tile(B, t + 1, vkeep, best, second);
        } else {
            rank_tile(A, t, vkeep, best, second);
        }
    } else {
        for (uint32_t t = 0; t < ntiles; ++t) {
            f32v4 acc[kGroups];
            score_tile<FP16, NKB, ADD_K>(acc, load_ops<NKB>(lds_table, lds_norm, t, lane, h), K, x);
            rank_tile(acc, t, vkeep, best, second);
        }
    }
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        best[g] = key_to_centroid(best[g], kmask, h);
        second[g] = key_to_centroid(second[g], kmask, h);
    }
}

// Order this wave's LDS accesses (lanes of one wave exchange data through
// LDS below; a wave's LDS operations complete in issue order, so only the
// compiler has to be kept from moving them -- no wait on global memory).
__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Number of lanes below this one whose bit is set in m (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Stage the 16-bit operand table and the norms in LDS (whole work-group);
// lds_normk (optional): the norms plus K, RN(||c||^2 + K), the same values
// the per-wave accumulator initialisation computes.
template <int NKB>
__device__ __forceinline__ void stage_table(u16x8 *lds_table, float *lds_norm,
                                            const u16x8 *__restrict__ table,
                                            const float *__restrict__ norm, uint32_t ntiles,
                                            float *lds_normk = nullptr, float K = 0.0f)
{
    const uint64_t entries = (uint64_t)ntiles * NKB * 64;
    for (uint64_t e = threadIdx.x; e < entries; e += blockDim.x)
        lds_table[e] = table[e];
    for (uint32_t c = threadIdx.x; c < ntiles * kTile; c += blockDim.x) {
        const float v = norm[c];
        lds_norm[c] = v;
        if (lds_normk)
            lds_normk[c] = v + K;
    }
    __syncthreads();
}

// Bound parameters for a job of dimension d and k centroids (host side).
inline MfmaParams make_params(bool fp16, uint64_t d, uint64_t k, bool subnormals_kept)
{
    MfmaParams prm;
    const double u = std::ldexp(1.0, -24);
    const uint64_t nkb = (d + kKBlock - 1) / kKBlock;
    const uint64_t ntiles = (k + kTile - 1) / kTile;
    const double P = (double)nkb * kKBlock;
    // a flushed input subnormal is off by at most the smallest normal
    prm.mu = subnormals_kept ? 0.0f : (fp16 ? (float)std::ldexp(1.0, -14) : (float)std::ldexp(1.0, -126));
    prm.gamma = (float)(2.0 * (P + 1.0) * u / (1.0 - 2.0 * (P + 1.0) * u));
    prm.key_bits = 1;
    while ((1ull << prm.key_bits) < ntiles * kTile)
        ++prm.key_bits;
    prm.trunc = (float)std::ldexp(1.0, (int)prm.key_bits - 23);
    prm.eps = (float)((double)(d + 3) * u * 1.001);
    prm.sqrt_d = (float)(std::sqrt((double)d) * (1.0 + 1e-6));
    return prm;
}

// Margin test of group g's point, after merging the (best, second) keys of
// the 4 lanes of its column (all lanes of the wave must call it).  unique:
// the best key's centroid is the exact argmin; otherwise every centroid that
// can be the argmin has trunc(score) <= thr (full: bounds unusable, every
// centroid is a candidate).
struct MarginDecision {
    uint32_t label;
    bool unique, full;
    float thr;
};

__device__ __forceinline__ MarginDecision decide(uint32_t best, uint32_t second, const WaveBounds &w,
                                          int g, uint32_t k, uint32_t kmask, bool table_ok,
                                          const MfmaParams &prm)
{
    const float tr1 = prm.trunc / (1.0f - prm.trunc), ep1 = prm.eps / (1.0f - prm.eps);
    uint32_t b = best, s = second;
    merge_pair(b, s, xor_swizzle<16>(b), xor_swizzle<16>(s));
    merge_pair(b, s, __shfl_xor(b, 32), __shfl_xor(s, 32));
    const float vb = __uint_as_float(b & ~kmask), vs = __uint_as_float(s & ~kmask);
    const float E = w.E0[g] + w.beta * w.K;
    const float ta = 1.01f * (2.0f * E + tr1 * vb + 2.0f * ep1 * (vb * (1.0f + tr1) + E)) + 1e-30f;
    MarginDecision v;
    // written so that NaN fails every "finite" test
    v.full = !(table_ok && vb >= 0.0f && vb + ta < 1e37f && w.K < 1e37f && w.Xn[g] < 1e18f);
    v.unique = !v.full && (b & kmask) < k && vs - vb > ta;
    v.label = b & kmask;
    v.thr = (vb + ta) * (1.0f + 1e-6f);
    return v;
}

// Exact FP32 distance (the reference's sequential fma from 0,
// lloyd_fused_feature_sum.cl:210-239) from the feature-major point and
// centroid arrays in global memory, loads issued 8 features ahead.
__device__ __forceinline__ float exact_dist_global(const float *__restrict__ pts, uint64_t n,
                                                   uint64_t p, const float *__restrict__ cen,
                                                   uint32_t d, uint32_t k, uint32_t c)
{
    const gfloat1 *gp = (const gfloat1 *)pts, *gc = (const gfloat1 *)cen;  // global loads
    float dist = 0.0f;
    for (uint32_t f0 = 0; f0 < d; f0 += 8) {
        float xv[8], cv[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t f = f0 + j < d ? f0 + j : 0;
            xv[j] = gp[(uint64_t)f * n + p];
            cv[j] = gc[(uint64_t)f * k + c];
        }
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j)
            if (f0 + j < d) {
                const float t = xv[j] - cv[j];
                dist = __builtin_fmaf(t, t, dist);
            }
    }
    return dist;
}

// One work-group builds a job's 16-bit operand table, norms and bounds from
// centroids [d][k] (the layout and stats[0..3] of mfma_prepare_kernel; see
// clk_mfma.hip).  red: blockDim.x doubles of LDS.
template <bool FP16>
__device__ void build_table_block(const float *__restrict__ cen, uint32_t d, uint32_t k,
                                  uint32_t ntiles, uint32_t nkb, u16x8 *__restrict__ table,
                                  float *__restrict__ norm, float *__restrict__ stats, double *red)
{
    const uint32_t tid = threadIdx.x;
    const uint64_t entries = (uint64_t)ntiles * nkb * 64;
    for (uint64_t e = tid; e < entries; e += blockDim.x) {
        const uint32_t lane = (uint32_t)(e % 64);
        const uint64_t tk = e / 64;
        const uint32_t kb = (uint32_t)(tk % nkb), t = (uint32_t)(tk / nkb);
        const uint32_t c = t * kTile + lane % 16;
        u16x8 v;
        for (int j = 0; j < 8; ++j) {
            const uint32_t f = kb * kKBlock + 8 * (lane / 16) + j;
            float x = (c < k && f < d) ? -2.0f * cen[(uint64_t)f * k + c] : 0.0f;
            if (!isfinite(x))
                x = 0.0f;
            v[j] = to16<FP16>(x);
        }
        table[e] = v;
    }
    double nmax = 0.0, rmax = 0.0;
    bool dropped = false;
    for (uint32_t c = tid; c < ntiles * kTile; c += blockDim.x) {
        double s = 0.0, r = 0.0;
        bool fin = c < k, in_range = true;
        for (uint32_t f = 0; f < d && fin; ++f) {
            const float v = cen[(uint64_t)f * k + c];
            fin = isfinite(v);
            in_range = in_range && (!FP16 || fabs(2.0 * (double)v) < 65504.0);
            s = fma((double)v, (double)v, s);
            const double res = (double)from16<FP16>(to16<FP16>(-2.0f * v)) + 2.0 * (double)v;
            r = fma(res, res, r);
        }
        const bool ok = fin && in_range && s < 3.0e38;
        dropped = dropped || (fin && !ok);
        norm[c] = ok ? (float)s : INFINITY;
        if (ok) {
            nmax = fmax(nmax, s);
            rmax = fmax(rmax, r);
        }
    }
    dropped = __syncthreads_or(dropped);
    double vals[2] = {nmax, rmax};
    for (int v = 0; v < 2; ++v) {
        red[tid] = vals[v];
        __syncthreads();
        uint32_t p2 = 1;
        while (p2 * 2 <= blockDim.x)
            p2 *= 2;
        if (tid < blockDim.x - p2)
            red[tid] = fmax(red[tid], red[tid + p2]);
        __syncthreads();
        for (uint32_t st = p2 / 2; st > 0; st /= 2) {
            if (tid < st)
                red[tid] = fmax(red[tid], red[tid + st]);
            __syncthreads();
        }
        vals[v] = red[0];
        __syncthreads();
    }
    if (tid == 0) {
        const double n = vals[0] * (1.0 + 1e-12);
        stats[0] = (float)(sqrt(n) * (1.0 + 1e-6));
        stats[1] = (float)(n * (1.0 + 1e-6));
        stats[2] = dropped ? 1.0f : 0.0f;
        stats[3] = (float)(sqrt(vals[1]) * (1.0 + 1e-6) + 1e-38);
    }
}

}  // namespace mfma
}  // namespace clk
