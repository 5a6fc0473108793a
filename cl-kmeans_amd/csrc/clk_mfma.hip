// Dense-distance precision form of the labeling pass: x·cᵀ on gfx950 MFMA
// (v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulation) plus an exact FP32
// relabel of every point whose best-to-second margin lies inside a derived
// error bound -- so labels equal the plain path's (sequential-fma distance,
// strict '<' scan from FLT_MAX, lowest index on ties) bit for bit.
//
// Scores.  S_c = ||c||^2 + K - 2 x·c = ||x - c||^2 - ||x||^2 + K.  The MFMA
// accumulator starts at RN(RN(||c||^2) + K); the operands are x and -2c
// rounded to 16 bits (x̃, c̃).  K >= ||x||^2 + E for every point of the wave (E the
// score error bound below), so every computed score is >= 0 and a float's bit
// pattern orders like its value: key = (bits(S) with the low KB bits replaced
// by c) orders by (truncated score, index); two keys a, b at a time update
// best = min3(best, a, b), second = min(second, med3(best, a, b)).
//
// Error bound of one computed score (u = 2^-24, μ = absolute input error of a
// 16-bit subnormal the matrix core might flush, P = padded dimension,
// X >= ||x||, C >= max ||c||, N >= max ||c||^2, and the measured rounding
// residuals  Rx >= ||x̃ - x|| (per point, from the converted values) and
// R >= max over centroids of ||c̃ - (-2c)||):
//   inputs           (X + Rx') (R + μ√d) + Rx' 2C,  Rx' = Rx + μ√d
//                    (Σ x̃(c̃ + 2c) + Σ (x̃ - x)(-2c), each by Cauchy-Schwarz)
//   MFMA summation   γ (N + K + Pm),  γ = 2(P+1)u / (1 - 2(P+1)u)  (each of at
//                    most P+1 additions may round or truncate), Pm the bound
//                    on Σ |x̃ c̃|
//   accumulator init u (2N + K)
// E = 1.01 × (sum) + (P+2) 2^-125 (flushed FP32 denormals).
//
// Margin.  With vb, vs the truncated best / second scores, τr = 2^(KB-23) the
// relative truncation, ε = (d+3)u the bound on |e_c - ||x-c||²| / ||x-c||² of
// the exact FP32 distance e_c, every centroid c != best satisfies
// e_c > e_best whenever  trunc(S_c) - vb > tau,
//   tau = 1.01 (2E + τr' vb + 2ε' (vb (1 + τr') + E)) + 1e-30.
// If vs - vb > tau the best key's centroid is the unique exact argmin.
//
// Refinement.  Only centroids with trunc(S_c) <= vb + tau can be the exact
// argmin; for them ||x - c||² - ||x||² <= T = (vb + tau)/(1 - τr) + E - K.
// The streaming kernel writes the best key's label and appends every other
// point, with T, to a queue.  A second kernel takes the queued points in
// dense batches of 64, scores them with its own K' and E', and evaluates the
// exact FP32 distance of every (point, centroid) pair with
// trunc(S'_c) <= T + K' + E' (a superset of those centroids), one pair per
// lane; the smallest distance below FLT_MAX, lowest index first, is the label
// (0 if none) -- the reference's strict-'<' scan restricted to the only
// centroids that can win.  A non-finite bound, score or sum makes T = +inf:
// every centroid is a candidate.
#include "clk_common.hpp"
#include "clk_kernels.hpp"

#include <algorithm>
#include <cmath>

namespace clk {
namespace {

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
__device__ __forceinline__ f32v4 mfma(u16x8 a, u16x8 b, f32v4 c)
{
    if constexpr (FP16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Whether the matrix cores keep 16-bit subnormal inputs: one wave multiplies
// 2^-20 (FP16 subnormal) and 2^-130 (BF16 subnormal) by 1.0 over K = 32;
// out[0] / out[1] = 1 when the exact products come back (the error bound then
// needs no flush term).
__global__ void __launch_bounds__(64) mfma_subnormal_check_kernel(float *out)
{
    u16x8 a16, a_bf, one16, one_bf;
    for (int j = 0; j < 8; ++j) {
        a16[j] = to16<true>(0x1p-20f);
        a_bf[j] = to16<false>(0x1p-130f);
        one16[j] = to16<true>(1.0f);
        one_bf[j] = to16<false>(1.0f);
    }
    const f32v4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    const f32v4 r16 = mfma<true>(a16, one16, z);
    const f32v4 rbf = mfma<false>(a_bf, one_bf, z);
    if (threadIdx.x == 0) {
        out[0] = r16[0] == 0x1p-15f ? 1.0f : 0.0f;
        out[1] = rbf[0] == 0x1p-125f ? 1.0f : 0.0f;
    }
}

// Per-iteration centroid tables (three small launches: clear, build, finish):
//   table[t][kb][lane] = 8 x 16-bit of -2c for centroid t*16 + lane%16,
//                        features kb*32 + 8*(lane/16) + j (0 past d or k)
//   norm[c] = RN(||c||^2) (FP32); +inf for a padded or non-finite centroid,
//             one whose ||c||^2 overflows, or (FP16) whose -2c overflows
//   cenT[c][f] = the FP32 centroids transposed (ntiles*16 rows of nkb*32, 0
//             past d or k), for the exact distances of the refinement
//   stats[0] = C >= max ||c||, stats[1] = N >= max ||c||^2, stats[3] = R >=
//   max ||c̃ + 2c|| over the others; stats[2] = 1 when a centroid with finite
//   coordinates had to be left out (its 16-bit or squared norm overflows):
//   then every point is scanned exactly; stats[4] (as uint32) = refinement
//   queue length; stats[12..15] = the build's running maxima (FP64 bits of
//   non-negative values order like the values) and the dropped flag.
__global__ void __launch_bounds__(64) mfma_prepare_clear_kernel(float *__restrict__ stats)
{
    if (threadIdx.x < 16)
        stats[threadIdx.x] = 0.0f;
}

template <bool FP16>
__global__ void __launch_bounds__(256)
mfma_prepare_kernel(const float *__restrict__ cen, uint32_t d, uint32_t k, uint32_t ntiles,
                    uint32_t nkb, u16x8 *__restrict__ table, float *__restrict__ norm,
                    float *__restrict__ stats, float *__restrict__ cenT)
{
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t entries = (uint64_t)ntiles * nkb * 64;
    for (uint64_t e = tid; e < entries; e += stride) {
        const uint32_t lane = (uint32_t)(e % 64);
        const uint64_t tk = e / 64;
        const uint32_t kb = (uint32_t)(tk % nkb), t = (uint32_t)(tk / nkb);
        const uint32_t c = t * kTile + lane % 16;
        u16x8 v;
        for (int j = 0; j < 8; ++j) {
            const uint32_t f = kb * kKBlock + 8 * (lane / 16) + j;
            float x = (c < k && f < d) ? -2.0f * cen[(uint64_t)f * k + c] : 0.0f;
            if (!isfinite(x))
                x = 0.0f;  // such a centroid is excluded through norm = +inf
            v[j] = to16<FP16>(x);
        }
        table[e] = v;
    }
    const uint32_t P = nkb * kKBlock;
    for (uint64_t e = tid; e < (uint64_t)ntiles * kTile * P; e += stride) {
        const uint32_t c = (uint32_t)(e / P), f = (uint32_t)(e % P);
        cenT[e] = (c < k && f < d) ? cen[(uint64_t)f * k + c] : 0.0f;
    }
    double nmax = 0.0, rmax = 0.0;
    bool dropped = false;  // a centroid with finite coordinates left out
    for (uint64_t c = tid; c < (uint64_t)ntiles * kTile; c += stride) {
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
    unsigned long long *acc = reinterpret_cast<unsigned long long *>(stats + 12);
    if (nmax > 0.0)
        atomicMax(&acc[0], (unsigned long long)__double_as_longlong(nmax));
    if (rmax > 0.0)
        atomicMax(&acc[1], (unsigned long long)__double_as_longlong(rmax));
    if (dropped)
        atomicOr(reinterpret_cast<uint32_t *>(stats) + 2, 1u);
}

__global__ void __launch_bounds__(64) mfma_prepare_finish_kernel(float *__restrict__ stats)
{
    if (threadIdx.x == 0) {
        const unsigned long long *acc = reinterpret_cast<const unsigned long long *>(stats + 12);
        const double n = __longlong_as_double((long long)acc[0]) * (1.0 + 1e-12);
        const double r = __longlong_as_double((long long)acc[1]);
        stats[0] = (float)(sqrt(n) * (1.0 + 1e-6));
        stats[1] = (float)(n * (1.0 + 1e-6));
        stats[2] = reinterpret_cast<const uint32_t *>(stats)[2] ? 1.0f : 0.0f;  // 1: exact scan
        stats[3] = (float)(sqrt(r) * (1.0 + 1e-6) + 1e-38);
    }
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
// (score & mask) | index stays one v_and_or_b32.
__device__ __forceinline__ uint32_t opaque(uint32_t v)
{
    asm volatile("" : "+v"(v));
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
template <bool FP16, int NKB>
__device__ __forceinline__ void score_tile(f32v4 (&acc)[kGroups], const TileOps<NKB> &o, float K,
                                           const PointOps<NKB> &x)
{
#pragma unroll
    for (int g = 0; g < kGroups; ++g)
        acc[g] = o.nrm + K;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
        for (int g = 0; g < kGroups; ++g)
            acc[g] = mfma<FP16>(o.a[kb], x.b[g][kb], acc[g]);
}

// Keys of one scored tile into the running (best, second) of each group.
__device__ __forceinline__ void rank_tile(const f32v4 (&acc)[kGroups], uint32_t c0, uint32_t kmask,
                                          uint32_t (&best)[kGroups], uint32_t (&second)[kGroups])
{
    uint32_t idx[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        idx[i] = opaque(c0 + i);
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        uint32_t key[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            key[i] = (__float_as_uint(acc[g][i]) & ~kmask) | idx[i];
        push2(best[g], second[g], key[0], key[1]);
        push2(best[g], second[g], key[2], key[3]);
    }
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
// X2[g] >= ||x||^2 and RX2[g] >= ||x̃ - x||^2 (FP32 sums over the 4 lanes of
// a point: at most P/4 + 2 roundings, plus underflow; x - x̃ is exact).
template <bool FP16, int NKB, bool FULL>
__device__ __forceinline__ void load_points(PointOps<NKB> &x, float (&X2)[kGroups],
                                            float (&RX2)[kGroups],
                                            const float *__restrict__ pts, uint64_t n, uint32_t d,
                                            const uint64_t (&pid)[kGroups], uint32_t q, uint32_t h,
                                            float *xs = nullptr)
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
        float s = 0.0f, r = 0.0f;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            u16x8 vb;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float v = raw[kb * 8 + j];
                s = __builtin_fmaf(v, v, s);
                vb[j] = to16<FP16>(v);
                const float res = v - from16<FP16>(vb[j]);
                r = __builtin_fmaf(res, res, r);
            }
            x.b[g][kb] = vb;
        }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        r += __shfl_xor(r, 16);
        r += __shfl_xor(r, 32);
        X2[g] = s * (1.0f + (float)(2 * NKB * kKBlock + 8) * kU) + 1e-37f;
        RX2[g] = r * (1.0f + (float)(2 * NKB * kKBlock + 8) * kU) + 1e-37f;
    }
}

// Everything the margin test needs about one wave's points.
struct WaveBounds {
    float K;               // shared score offset
    float beta;            // K-proportional part of the score error
    float E0[kGroups];     // the rest of the score error bound, per point
    float Xn[kGroups];     // >= ||x||
};

template <int NKB>
__device__ __forceinline__ WaveBounds wave_bounds(const float (&X2)[kGroups],
                                                  const float (&RX2)[kGroups], const float *stats,
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
        const float Rx = sqrtf(RX2[g]) * (1.0f + 4.0f * kU) + ms;
        const float in = (X + Rx) * (R + ms) + Rx * 2.0f * Cx;
        const float Pm = (X + Rx) * (2.0f * Cx + R + ms);
        w.Xn[g] = X;
        w.E0[g] = 1.01f * (in + gm * (Nx + Pm) + 2.0f * kU * Nx) + absP;
        kneed = fmaxf(kneed, X2[g] + w.E0[g]);  // a NaN point does not poison K
    }
    for (int o = 1; o < 64; o <<= 1)
        kneed = fmaxf(kneed, __shfl_xor(kneed, o));
    w.beta = 1.01f * (gm + kU);
    w.K = kneed * (1.0f + 1e-5f) / (1.0f - w.beta) + 1e-30f;
    return w;
}

// Best / second keys over all centroid tiles, software-pipelined: the MFMAs
// of tile t run while the keys of tile t-1 are ranked, and the LDS operands
// of the next tile are already in flight (two accumulator and two operand
// sets in turn; at the loop head accA holds the scored, unranked tile t-1
// and o1 the operands of tile t).
template <bool FP16, int NKB>
__device__ __forceinline__ void rank_all(const u16x8 *lds_table, const float *lds_norm,
                                         uint32_t ntiles, uint32_t lane, uint32_t h, float K,
                                         uint32_t kmask, const PointOps<NKB> &x,
                                         uint32_t (&best)[kGroups], uint32_t (&second)[kGroups])
{
#pragma unroll
    for (int g = 0; g < kGroups; ++g)
        best[g] = second[g] = 0xFFFFFFFFu;
    const uint32_t last = ntiles - 1;
    f32v4 accA[kGroups], accB[kGroups];
    TileOps<NKB> o0 = load_ops<NKB>(lds_table, lds_norm, 0, lane, h);
    TileOps<NKB> o1 = load_ops<NKB>(lds_table, lds_norm, min(1u, last), lane, h);
    score_tile<FP16, NKB>(accA, o0, K, x);
    uint32_t t = 1;
    for (; t + 1 < ntiles; t += 2) {
        o0 = load_ops<NKB>(lds_table, lds_norm, t + 1, lane, h);
        score_tile<FP16, NKB>(accB, o1, K, x);
        rank_tile(accA, (t - 1) * kTile + 4 * h, kmask, best, second);
        o1 = load_ops<NKB>(lds_table, lds_norm, min(t + 2, last), lane, h);
        score_tile<FP16, NKB>(accA, o0, K, x);
        rank_tile(accB, t * kTile + 4 * h, kmask, best, second);
    }
    if (t < ntiles) {
        score_tile<FP16, NKB>(accB, o1, K, x);
        rank_tile(accA, (t - 1) * kTile + 4 * h, kmask, best, second);
        rank_tile(accB, t * kTile + 4 * h, kmask, best, second);
    } else {
        rank_tile(accA, (t - 1) * kTile + 4 * h, kmask, best, second);
    }
}

// Order this wave's LDS accesses (lanes of one wave exchange data through
// LDS below; a wave's LDS operations complete in issue order, so only the
// compiler has to be kept from moving them -- no wait on global memory).
__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt(uint32_t lane)
{
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// Stage the 16-bit operand table and the norms in LDS (whole work-group).
template <int NKB>
__device__ __forceinline__ void stage_table(u16x8 *lds_table, float *lds_norm,
                                            const u16x8 *__restrict__ table,
                                            const float *__restrict__ norm, uint32_t ntiles)
{
    const uint64_t entries = (uint64_t)ntiles * NKB * 64;
    for (uint64_t e = threadIdx.x; e < entries; e += blockDim.x)
        lds_table[e] = table[e];
    for (uint32_t c = threadIdx.x; c < ntiles * kTile; c += blockDim.x)
        lds_norm[c] = norm[c];
    __syncthreads();
}

// Streaming kernel: work-group of fast_waves(NKB) waves (up to d = 64 two per
// SIMD sharing one LDS table, so each wave's point loads hide behind the
// other's MFMA work; one per SIMD above, for the registers), each wave labels
// 64 consecutive points per step.  Points whose margin test fails are
// appended to qidx[] / qthr[] (T, above; count at *qlen).
constexpr int fast_waves(int nkb) { return nkb <= 2 ? 8 : 4; }

template <bool FP16, int NKB, bool FULL>
__global__ void __launch_bounds__(64 * fast_waves(NKB))
label_mfma_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                  const u16x8 *__restrict__ table, const float *__restrict__ norm,
                  const float *__restrict__ stats, uint32_t ntiles, MfmaParams prm,
                  uint32_t *__restrict__ labels, uint32_t *__restrict__ qidx,
                  float *__restrict__ qthr, uint32_t *__restrict__ qlen)
{
    extern __shared__ __align__(16) u16x8 lds_table[];
    float *lds_norm = reinterpret_cast<float *>(lds_table + (uint64_t)ntiles * NKB * 64);
    stage_table<NKB>(lds_table, lds_norm, table, norm, ntiles);
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const uint32_t q = lane % 16, h = lane / 16;
    const bool table_ok = stats[2] == 0.0f;
    const uint32_t kmask = (1u << prm.key_bits) - 1u;
    const float tr1 = prm.trunc / (1.0f - prm.trunc), ep1 = prm.eps / (1.0f - prm.eps);

    constexpr int kWaves = fast_waves(NKB);
    const uint64_t batches = (n + 63) / 64;
    for (uint64_t bt = (uint64_t)blockIdx.x * kWaves + wave; bt < batches;
         bt += (uint64_t)gridDim.x * kWaves) {
        uint64_t pid[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g)
            pid[g] = bt * 64 + g * 16 + q;
        PointOps<NKB> x;
        float X2[kGroups], RX2[kGroups];
        load_points<FP16, NKB, FULL>(x, X2, RX2, pts, n, d, pid, q, h);
        const WaveBounds w = wave_bounds<NKB>(X2, RX2, stats, prm);
        uint32_t best[kGroups], second[kGroups];
        rank_all<FP16, NKB>(lds_table, lds_norm, ntiles, lane, h, w.K, kmask, x, best, second);

        // margin test; lane l owns point bt*64 + l (group h, column q)
        uint32_t mine = 0;
        bool amb = false;
        float T = 0.0f;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            uint32_t b = best[g], s = second[g];
            merge_pair(b, s, __shfl_xor(b, 16), __shfl_xor(s, 16));
            merge_pair(b, s, __shfl_xor(b, 32), __shfl_xor(s, 32));
            const float vb = __uint_as_float(b & ~kmask), vs = __uint_as_float(s & ~kmask);
            const float E = w.E0[g] + w.beta * w.K;
            const float ta = 1.01f * (2.0f * E + tr1 * vb + 2.0f * ep1 * (vb * (1.0f + tr1) + E)) + 1e-30f;
            // written so that NaN fails every "finite" test
            const bool full = !(table_ok && vb >= 0.0f && vb + ta < 1e37f && w.K < 1e37f && w.Xn[g] < 1e18f);
            const bool unique = !full && (b & kmask) < k && vs - vb > ta;
            const float tv = (vb + ta) * (1.0f + tr1) + E - w.K;
            const float Tg = full ? INFINITY : tv + 4.0f * kU * (vb + ta + E + w.K);
            if ((uint32_t)g == h) {
                mine = b & kmask;
                amb = !unique;
                T = Tg;
            }
        }
        const uint64_t p = bt * 64 + lane;
        if (p < n)
            labels[p] = mine;  // provisional for queued points
        const uint64_t m = __ballot(amb && p < n);
        if (m != 0) {
            const uint32_t first = (uint32_t)__builtin_ctzll(m);
            uint32_t slot0 = 0;
            if (lane == first)
                slot0 = atomicAdd(qlen, (uint32_t)__builtin_popcountll(m));
            slot0 = __shfl(slot0, (int)first);
            if (amb && p < n) {
                const uint32_t slot = slot0 + (uint32_t)__builtin_popcountll(m & lanemask_lt(lane));
                qidx[slot] = (uint32_t)p;
                qthr[slot] = T;
            }
        }
    }
}

constexpr uint32_t kPairCap = 1024;  // candidate pairs per wave and flush
constexpr int kRefineWaves = 4;

// Refinement kernel for the queued points (see the header).  The 16-bit table
// and norms are read from L2 (tile t+1 prefetched while tile t is scored);
// LDS holds, per wave, the 64 points' FP32 rows, a pair list, the point ids,
// the running (distance bits << 32 | centroid) minimum of each point and a
// pair counter.  cenT = the centroids transposed, [ntiles*16][P] FP32.
template <bool FP16, int NKB, bool FULL>
__global__ void __launch_bounds__(64 * kRefineWaves)
refine_mfma_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                   const float *__restrict__ cenT, const u16x8 *__restrict__ table,
                   const float *__restrict__ norm, const float *__restrict__ stats,
                   uint32_t ntiles, MfmaParams prm, uint32_t *__restrict__ labels,
                   const uint32_t *__restrict__ qidx, const float *__restrict__ qthr,
                   const uint32_t *__restrict__ qlen)
{
    constexpr uint32_t P = NKB * kKBlock;
    const uint32_t qn = *qlen;
    const uint64_t batches = ((uint64_t)qn + 63) / 64;
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    extern __shared__ __align__(16) float lds_x[];
    float *xs = lds_x + wave * 64 * P;
    uint32_t *pairs = reinterpret_cast<uint32_t *>(lds_x + kRefineWaves * 64 * P) + wave * kPairCap;
    unsigned long long *lds_min = reinterpret_cast<unsigned long long *>(
        reinterpret_cast<uint32_t *>(lds_x + kRefineWaves * 64 * P) + kRefineWaves * kPairCap);
    unsigned long long *wmin = lds_min + wave * 64;
    uint32_t *wcount = reinterpret_cast<uint32_t *>(lds_min + kRefineWaves * 64) + wave;
    const uint32_t q = lane % 16, h = lane / 16;
    const uint32_t kmask = (1u << prm.key_bits) - 1u;

    for (uint64_t bt = (uint64_t)blockIdx.x * kRefineWaves + wave; bt < batches;
         bt += (uint64_t)gridDim.x * kRefineWaves) {
        uint64_t pid[kGroups];
        float thr_in[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const uint64_t slot = bt * 64 + g * 16 + q;
            const bool ok = slot < qn;
            pid[g] = ok ? (uint64_t)qidx[slot] : ~0ull;
            thr_in[g] = ok ? qthr[slot] : 0.0f;
        }
        PointOps<NKB> x;
        float X2[kGroups], RX2[kGroups];
        load_points<FP16, NKB, FULL>(x, X2, RX2, pts, n, d, pid, q, h, xs);
        const WaveBounds w = wave_bounds<NKB>(X2, RX2, stats, prm);
        // candidate threshold on this kernel's scores: T + K' + E', rounded up;
        // NaN or infinite -> every centroid is a candidate
        float thr[kGroups];
        bool all[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const float E = w.E0[g] + w.beta * w.K;
            const float tv = thr_in[g] + w.K + E;
            thr[g] = tv + 4.0f * kU * (fabsf(thr_in[g]) + w.K + E);
            all[g] = !(thr[g] < 1e37f && w.K < 1e37f && w.Xn[g] < 1e18f);
        }
        uint64_t mypid = pid[0];
#pragma unroll
        for (int g = 1; g < kGroups; ++g)
            if ((uint32_t)g == h)
                mypid = pid[g];
        wmin[lane] = ~0ull;  // slot lane = 16 h + q
        if (lane == 0)
            *wcount = 0;
        wave_lds_order();

        // Exact distances of the collected pairs, the running minimum of
        // (distance bits, centroid) per point; order does not matter.
        auto flush = [&](uint32_t count) {
            for (uint32_t r = lane; r < count; r += 64) {
                const uint32_t pr = pairs[r];
                const uint32_t slot = pr >> 20, c = pr & 0xFFFFFu;
                const float e = exact_dist(xs + slot * P, cenT + (uint64_t)c * P, d);
                if (e < FLT_MAX)  // the reference's strict '<' from FLT_MAX
                    atomicMin(&wmin[slot], ((unsigned long long)__float_as_uint(e) << 32) | c);
            }
            wave_lds_order();
            if (lane == 0)
                *wcount = 0;
            wave_lds_order();
        };
        const uint32_t last = ntiles - 1;
        TileOps<NKB> nxt = load_ops<NKB>(table, norm, 0, lane, h);
        for (uint32_t t = 0; t < ntiles; ++t) {
            const TileOps<NKB> cur = nxt;
            nxt = load_ops<NKB>(table, norm, min(t + 1, last), lane, h);
            f32v4 acc[kGroups];
            score_tile<FP16, NKB>(acc, cur, w.K, x);
            uint32_t cmask = 0;  // bit 4g + i: (group g's point, centroid t*16 + 4h + i)
#pragma unroll
            for (int g = 0; g < kGroups; ++g)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t c = t * kTile + 4 * h + i;
                    const float sc = __uint_as_float(__float_as_uint(acc[g][i]) & ~kmask);
                    const bool cand = pid[g] < n && c < k && (all[g] || !(sc > thr[g]));
                    cmask |= (uint32_t)cand << (4 * g + i);
                }
            const uint32_t nl = (uint32_t)__builtin_popcountll(__ballot(cmask != 0));
            if (nl == 0)
                continue;
            // at most 16 pairs per such lane: flush first if they may not fit
            const uint32_t before = *wcount;
            wave_lds_order();
            if (before + 16 * nl > kPairCap)
                flush(min(before, kPairCap));
            while (cmask != 0) {
                const uint32_t bit = (uint32_t)__builtin_ctz(cmask);
                cmask &= cmask - 1;
                const uint32_t g = bit / 4, i = bit % 4;
                const uint32_t at = atomicAdd(wcount, 1u);
                pairs[at] = ((g * 16 + q) << 20) | (t * kTile + 4 * h + i);
            }
            wave_lds_order();
        }
        flush(min(*wcount, kPairCap));
        const unsigned long long mn = wmin[lane];
        if (mypid < n)
            labels[mypid] = (mn == ~0ull) ? 0u : (uint32_t)mn;
        wave_lds_order();
    }
}

using LabelKernelM = void (*)(const float *, uint64_t, uint32_t, uint32_t, const u16x8 *,
                              const float *, const float *, uint32_t, MfmaParams, uint32_t *,
                              uint32_t *, float *, uint32_t *);
using RefineKernelM = void (*)(const float *, uint64_t, uint32_t, uint32_t, const float *,
                               const u16x8 *, const float *, const float *, uint32_t, MfmaParams,
                               uint32_t *, const uint32_t *, const float *, const uint32_t *);

template <bool FP16, bool FULL>
void pick_mfma(uint32_t nkb, LabelKernelM &lk, RefineKernelM &rk)
{
    switch (nkb) {
    case 1: lk = label_mfma_kernel<FP16, 1, FULL>; rk = refine_mfma_kernel<FP16, 1, FULL>; break;
    case 2: lk = label_mfma_kernel<FP16, 2, FULL>; rk = refine_mfma_kernel<FP16, 2, FULL>; break;
    case 3: lk = label_mfma_kernel<FP16, 3, FULL>; rk = refine_mfma_kernel<FP16, 3, FULL>; break;
    default: lk = label_mfma_kernel<FP16, 4, FULL>; rk = refine_mfma_kernel<FP16, 4, FULL>; break;
    }
}

struct MfmaLayout {
    uint32_t ntiles, nkb;
    size_t table_bytes, norm_bytes, cent_bytes, refine_lds;
};

MfmaLayout mfma_layout(uint64_t d, uint64_t k)
{
    MfmaLayout l;
    l.ntiles = (uint32_t)div_up(k, (uint64_t)kTile);
    l.nkb = (uint32_t)div_up(d, (uint64_t)kKBlock);
    l.table_bytes = (size_t)l.ntiles * l.nkb * 64 * sizeof(u16x8);
    l.norm_bytes = (size_t)l.ntiles * kTile * sizeof(float);
    l.cent_bytes = (size_t)l.ntiles * kTile * l.nkb * kKBlock * sizeof(float);
    l.refine_lds = kRefineWaves * (64 * l.nkb * kKBlock * sizeof(float) + kPairCap * sizeof(uint32_t) +
                                   64 * sizeof(unsigned long long) + sizeof(uint32_t));
    return l;
}

hipError_t allow_lds(const void *kern, size_t lds)
{
    if (lds <= 65536)
        return hipSuccess;
    return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

}  // namespace

bool mfma_labeling_supported(const clk_context *ctx, uint64_t d, uint64_t n, uint64_t k)
{
    if (d == 0 || d > (uint64_t)kMaxKb * kKBlock || k == 0 || k > (1u << 20) || n >= (1ull << 32))
        return false;
    const MfmaLayout l = mfma_layout(d, k);
    return l.table_bytes + l.norm_bytes <= ctx->lds_per_block && l.refine_lds <= ctx->lds_per_block;
}

int launch_labeling_mfma(clk_context *ctx, bool fp16, uint64_t d, uint64_t n, uint64_t k,
                         const float *points, const float *centroids, uint32_t *labels,
                         hipStream_t stream)
{
    CLK_REQUIRE(mfma_labeling_supported(ctx, d, n, k), CLK_ERR_UNSUPPORTED,
                "MFMA labeling: needs d <= 128, n < 2^32 and a 16-bit centroid table that fits in LDS");
    const MfmaLayout l = mfma_layout(d, k);
    CLK_TRY(ensure_scratch(ctx->mfma_tables, l.table_bytes + l.norm_bytes + 64 + l.cent_bytes));
    CLK_TRY(ensure_scratch(ctx->mfma_queue, std::max<uint64_t>(n, 1) * 2 * sizeof(uint32_t)));
    u16x8 *table = static_cast<u16x8 *>(ctx->mfma_tables.ptr);
    float *norm = reinterpret_cast<float *>(static_cast<char *>(ctx->mfma_tables.ptr) + l.table_bytes);
    float *stats = norm + (size_t)l.ntiles * kTile;  // 16 floats
    float *cenT = stats + 16;
    uint32_t *qlen = reinterpret_cast<uint32_t *>(stats + 4);
    uint32_t *qidx = static_cast<uint32_t *>(ctx->mfma_queue.ptr);
    float *qthr = reinterpret_cast<float *>(qidx + std::max<uint64_t>(n, 1));
    ctx->mfma_qlen = qlen;
    {
        KernelTimer t(ctx, "MfmaPrepare", stream);
        hipLaunchKernelGGL(mfma_prepare_clear_kernel, dim3(1), dim3(64), 0, stream, stats);
        const uint64_t work = std::max<uint64_t>((uint64_t)l.ntiles * kTile * l.nkb * kKBlock,
                                                 (uint64_t)l.ntiles * l.nkb * 64);
        const unsigned pgrid = (unsigned)std::min<uint64_t>(div_up(work, (uint64_t)256), 1024);
        if (fp16)
            hipLaunchKernelGGL(mfma_prepare_kernel<true>, dim3(pgrid), dim3(256), 0, stream, centroids,
                               (uint32_t)d, (uint32_t)k, l.ntiles, l.nkb, table, norm, stats, cenT);
        else
            hipLaunchKernelGGL(mfma_prepare_kernel<false>, dim3(pgrid), dim3(256), 0, stream, centroids,
                               (uint32_t)d, (uint32_t)k, l.ntiles, l.nkb, table, norm, stats, cenT);
        hipLaunchKernelGGL(mfma_prepare_finish_kernel, dim3(1), dim3(64), 0, stream, stats);
        CLK_HIP_CHECK(hipGetLastError());
    }
    if (n == 0)
        return CLK_OK;
    if (ctx->mfma_subnormals < 0) {  // once per context
        hipLaunchKernelGGL(mfma_subnormal_check_kernel, dim3(1), dim3(64), 0, stream, stats + 8);
        CLK_HIP_CHECK(hipGetLastError());
        float flags[2];
        CLK_HIP_CHECK(hipMemcpyAsync(flags, stats + 8, sizeof(flags), hipMemcpyDeviceToHost, stream));
        CLK_HIP_CHECK(hipStreamSynchronize(stream));
        ctx->mfma_subnormals = (flags[0] == 1.0f ? 1 : 0) | (flags[1] == 1.0f ? 2 : 0);
    }
    MfmaParams prm;
    const double u = std::ldexp(1.0, -24);
    const double P = (double)l.nkb * kKBlock;
    // a flushed input subnormal is off by at most the smallest normal
    const bool kept = (ctx->mfma_subnormals & (fp16 ? 1 : 2)) != 0;
    prm.mu = kept ? 0.0f : (fp16 ? (float)std::ldexp(1.0, -14) : (float)std::ldexp(1.0, -126));
    prm.gamma = (float)(2.0 * (P + 1.0) * u / (1.0 - 2.0 * (P + 1.0) * u));
    prm.key_bits = 1;
    while ((1ull << prm.key_bits) < (uint64_t)l.ntiles * kTile)
        ++prm.key_bits;
    prm.trunc = (float)std::ldexp(1.0, (int)prm.key_bits - 23);
    prm.eps = (float)((double)(d + 3) * u * 1.001);
    prm.sqrt_d = (float)(std::sqrt((double)d) * (1.0 + 1e-6));

    const bool full = d == (uint64_t)l.nkb * kKBlock;
    LabelKernelM lk;
    RefineKernelM rk;
    if (fp16)
        full ? pick_mfma<true, true>(l.nkb, lk, rk) : pick_mfma<true, false>(l.nkb, lk, rk);
    else
        full ? pick_mfma<false, true>(l.nkb, lk, rk) : pick_mfma<false, false>(l.nkb, lk, rk);
    const size_t lds = l.table_bytes + l.norm_bytes;
    const size_t lds_r = l.refine_lds;
    CLK_HIP_CHECK(allow_lds((const void *)lk, lds));
    CLK_HIP_CHECK(allow_lds((const void *)rk, lds_r));
    const int fast_block = 64 * fast_waves((int)l.nkb);
    const uint64_t grid = auto_grid(ctx, (const void *)lk, fast_block, lds, div_up(n, (uint64_t)64) * 64);
    {
        KernelTimer t(ctx, fp16 ? "LabelingMfmaF16" : "LabelingMfmaBF16", stream);
        hipLaunchKernelGGL(lk, dim3((unsigned)grid), dim3(fast_block), lds, stream, points, n,
                           (uint32_t)d, (uint32_t)k, table, norm, stats, l.ntiles, prm, labels,
                           qidx, qthr, qlen);
        CLK_HIP_CHECK(hipGetLastError());
    }
    // a resident grid; blocks past the queue return at once
    const int refine_block = 64 * kRefineWaves;
    const uint64_t grid_r = auto_grid(ctx, (const void *)rk, refine_block, lds_r, div_up(n, (uint64_t)64) * 64);
    KernelTimer t(ctx, "LabelingMfmaRefine", stream);
    hipLaunchKernelGGL(rk, dim3((unsigned)grid_r), dim3(refine_block), lds_r, stream, points, n,
                       (uint32_t)d, (uint32_t)k, cenT, table, norm, stats, l.ntiles, prm, labels,
                       qidx, qthr, qlen);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

}  // namespace clk

extern "C" int clk_labeling_refined_points(clk_context *ctx, uint64_t *points)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(points != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_labeling_refined_points: null out");
    *points = 0;
    if (!ctx->mfma_qlen)
        return CLK_OK;
    uint32_t v = 0;
    CLK_HIP_CHECK(hipDeviceSynchronize());
    CLK_HIP_CHECK(hipMemcpy(&v, ctx->mfma_qlen, sizeof(uint32_t), hipMemcpyDeviceToHost));
    *points = v;
    return CLK_OK;
}
