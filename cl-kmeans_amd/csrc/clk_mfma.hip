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
// X >= ||x||, C >= max ||c||, N >= max ||c||^2, the rounding bound
// Rx >= ||x̃ - x|| = u16 X + s16 √d (per point: relative 2^-11 / 2^-8 per
// coordinate for FP16 / BF16, half the 16-bit subnormal spacing below the
// normal range; mfma::rounding_bound) and the measured centroid residual
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
#include "clk_mfma_device.hpp"

#include <algorithm>
#include <cmath>

namespace clk {
namespace {

using namespace mfma;

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
    const f32v4 r16 = mma16<true>(a16, one16, z);
    const f32v4 rbf = mma16<false>(a_bf, one_bf, z);
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

// stats[5] = the launch-wide score offset K_L for this labeling: 1.02 x the
// largest per-wave K of the previous labeling on this context (hint[0],
// reset here for this labeling to record its own), or 0 (none).  A wave
// whose own K is <= K_L uses K_L with the norms pre-offset in LDS; any K >=
// ||x||^2 + E is valid, so K_L only changes which waves take that path.
__global__ void __launch_bounds__(64) mfma_prepare_finish_kernel(float *__restrict__ stats, float *__restrict__ hint)
{
    if (threadIdx.x == 0) {
        const float hv = hint[0];
        stats[5] = (hv > 0.0f && hv < 1e30f) ? hv * 1.02f : 0.0f;
        hint[0] = 0.0f;
        const unsigned long long *acc = reinterpret_cast<const unsigned long long *>(stats + 12);
        const double n = __longlong_as_double((long long)acc[0]) * (1.0 + 1e-12);
        const double r = __longlong_as_double((long long)acc[1]);
        stats[0] = (float)(sqrt(n) * (1.0 + 1e-6));
        stats[1] = (float)(n * (1.0 + 1e-6));
        stats[2] = reinterpret_cast<const uint32_t *>(stats)[2] ? 1.0f : 0.0f;  // 1: exact scan
        stats[3] = (float)(sqrt(r) * (1.0 + 1e-6) + 1e-38);
    }
}



// Streaming kernel: work-group of fast_waves(NKB) waves (up to d = 64 three
// per SIMD sharing one LDS table, so each wave's point loads and LDS reads
// hide behind the others' MFMA and ranking work, and each wave holds two
// tiles' scores so that its next MFMAs run under its ranking -- at d = 64,
// k = 1024 2-4% faster than four per SIMD with one tile's scores, which was 8%
// faster than two per SIMD; one per SIMD above, for the registers), each
// wave labels 64 consecutive points per step.  Points whose margin test fails are
// appended to qidx[] / qthr[] (T, above; count at *qlen).
constexpr int fast_waves(int nkb) { return nkb <= 2 ? 12 : 4; }

// Score offset: each wave needs K >= ||x||^2 + E over its points (wave_bounds).
// With a launch-wide K_L (stats[5]; 0 = none) the norms are staged a second
// time as RN(||c||^2 + K_L), and a wave whose own K is <= K_L starts its
// accumulators from those (no add per tile); the others add their own K.
// Every wave's own K goes to hint[0] (atomic max of positive float bits) for
// the next labeling's K_L.
template <bool FP16, int NKB, bool FULL>
__global__ void __launch_bounds__(64 * fast_waves(NKB))
label_mfma_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, uint32_t k,
                  const u16x8 *__restrict__ table, const float *__restrict__ norm,
                  const float *__restrict__ stats, uint32_t ntiles, MfmaParams prm,
                  uint32_t *__restrict__ labels, uint32_t *__restrict__ qidx,
                  float *__restrict__ qthr, uint32_t *__restrict__ qlen, float *__restrict__ hint,
                  const u16x8 *__restrict__ pin_ops, const f32v4 *__restrict__ pin_x2)
{
    extern __shared__ __align__(16) u16x8 lds_table[];
    float *lds_norm = reinterpret_cast<float *>(lds_table + (uint64_t)ntiles * NKB * 64);
    float *lds_normk = lds_norm + (size_t)ntiles * kTile;
    const float KL = stats[5];
    stage_table<NKB>(lds_table, lds_norm, table, norm, ntiles, KL > 0.0f ? lds_normk : nullptr, KL);
    float kmax = 0.0f;  // this wave's largest own K
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
        float X2[kGroups];
        if (pin_ops) {  // the pinned copy: contiguous 128-bit loads
#pragma unroll
            for (int g = 0; g < kGroups; ++g)
#pragma unroll
                for (int kb = 0; kb < NKB; ++kb)
                    x.b[g][kb] = pin_ops[((bt * kGroups + g) * NKB + kb) * 64 + lane];
            const f32v4 x2 = pin_x2[bt * 16 + q];
#pragma unroll
            for (int g = 0; g < kGroups; ++g)
                X2[g] = x2[g];
        } else {
            load_points<FP16, NKB, FULL>(x, X2, pts, n, d, pid, q, h);
        }
        WaveBounds w = wave_bounds<NKB, FP16>(X2, stats, prm);
        kmax = fmaxf(kmax, w.K);  // (a NaN K is skipped)
        uint32_t best[kGroups], second[kGroups];
        if (KL > 0.0f && w.K <= KL) {  // wave-uniform
            w.K = KL;
            rank_all<FP16, NKB, false, true>(lds_table, lds_normk, ntiles, lane, h, w.K, kmask, x, best, second);
        } else {
            rank_all<FP16, NKB, true, true>(lds_table, lds_norm, ntiles, lane, h, w.K, kmask, x, best, second);
        }

        // margin test; lane l owns point bt*64 + l (group h, column q)
        uint32_t mine = 0;
        bool amb = false;
        float T = 0.0f;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            uint32_t b = best[g], s = second[g];
            merge_pair(b, s, xor_swizzle<16>(b), xor_swizzle<16>(s));
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
                const uint32_t slot = slot0 + lanes_below(m);
                qidx[slot] = (uint32_t)p;
                qthr[slot] = T;
            }
        }
    }
    if (lane == 0 && kmax > 0.0f)
        atomicMax(reinterpret_cast<uint32_t *>(hint), __float_as_uint(kmax));
}

// The pinned operand copy of points[d][n] (clk_labeling_pin_points): for
// each 64-point step bt, group g and K-block kb, the 64 lanes' B operands in
// lane order -- ops[((bt * 4 + g) * NKB + kb) * 64 + lane] -- and each
// point's X2 at x2[bt * 16 + q][g]: exactly what load_points gives the
// labeling kernel's lane (q, h).  One wave per step.
template <bool FP16, int NKB>
__global__ void __launch_bounds__(256)
mfma_pin_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d, u16x8 *__restrict__ ops,
                float *__restrict__ x2)
{
    const uint32_t lane = threadIdx.x % 64, q = lane % 16, h = lane / 16;
    const uint64_t bt = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (bt >= (n + 63) / 64)
        return;
    uint64_t pid[kGroups];
#pragma unroll
    for (int g = 0; g < kGroups; ++g)
        pid[g] = bt * 64 + g * 16 + q;
    PointOps<NKB> x;
    float X2[kGroups];
    float unused[kGroups][NKB * 8];
    load_points_core<FP16, NKB, false, false>(x, X2, pts, n, d, pid, q, h, nullptr, unused);
#pragma unroll
    for (int g = 0; g < kGroups; ++g)
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb)
            ops[((bt * kGroups + g) * NKB + kb) * 64 + lane] = x.b[g][kb];
    // lane (q, h) writes group h's X2 (the four lanes of a column hold it)
    float mine = X2[0];
#pragma unroll
    for (int g = 1; g < kGroups; ++g)
        mine = (uint32_t)g == h ? X2[g] : mine;
    x2[(bt * 16 + q) * 4 + h] = mine;
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
        float X2[kGroups];
        load_points<FP16, NKB, FULL>(x, X2, pts, n, d, pid, q, h, xs);
        const WaveBounds w = wave_bounds<NKB, FP16>(X2, stats, prm);
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
                              uint32_t *, float *, uint32_t *, float *, const u16x8 *, const f32v4 *);
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

// Operands (u16x8) of the pinned copy of n points with nkb K-blocks; the
// X2 array (one f32v4 per 16 points) follows them.
uint64_t pin_ops_count(uint64_t n, uint32_t nkb) { return div_up(n, (uint64_t)64) * kGroups * nkb * 64; }
uint64_t pin_bytes(uint64_t n, uint32_t nkb)
{
    return pin_ops_count(n, nkb) * sizeof(u16x8) + div_up(n, (uint64_t)64) * 16 * sizeof(f32v4);
}

hipError_t allow_lds(const void *kern, size_t lds)
{
    if (lds <= 65536)
        return hipSuccess;
    return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

}  // namespace

int mfma_subnormal_flags(clk_context *ctx, hipStream_t stream)
{
    if (ctx->mfma_subnormals >= 0)
        return CLK_OK;
    float *flags_d = nullptr;
    CLK_HIP_CHECK(hipMalloc(&flags_d, 2 * sizeof(float)));
    hipLaunchKernelGGL(mfma_subnormal_check_kernel, dim3(1), dim3(64), 0, stream, flags_d);
    hipError_t e = hipGetLastError();
    float flags[2] = {0.0f, 0.0f};
    if (e == hipSuccess)
        e = hipMemcpyAsync(flags, flags_d, sizeof(flags), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    (void)hipFree(flags_d);
    CLK_HIP_CHECK(e);
    ctx->mfma_subnormals = (flags[0] == 1.0f ? 1 : 0) | (flags[1] == 1.0f ? 2 : 0);
    return CLK_OK;
}

bool mfma_labeling_supported(const clk_context *ctx, uint64_t d, uint64_t n, uint64_t k)
{
    if (d == 0 || d > (uint64_t)kMaxKb * kKBlock || k == 0 || k > (1u << 20) || n >= (1ull << 32))
        return false;
    const MfmaLayout l = mfma_layout(d, k);
    return l.table_bytes + 2 * l.norm_bytes <= ctx->lds_per_block && l.refine_lds <= ctx->lds_per_block;
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
    if (!ctx->mfma_hint.ptr) {  // no offset hint before the first labeling
        CLK_TRY(ensure_scratch(ctx->mfma_hint, 256));
        CLK_TRY(clk_fill_zero(ctx, ctx->mfma_hint.ptr, 256, stream));
    }
    float *hint = static_cast<float *>(ctx->mfma_hint.ptr);
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
        hipLaunchKernelGGL(mfma_prepare_finish_kernel, dim3(1), dim3(64), 0, stream, stats, hint);
        CLK_HIP_CHECK(hipGetLastError());
    }
    if (n == 0)
        return CLK_OK;
    CLK_TRY(mfma_subnormal_flags(ctx, stream));
    const MfmaParams prm = make_params(fp16, d, k, (ctx->mfma_subnormals & (fp16 ? 1 : 2)) != 0);

    const bool full = d == (uint64_t)l.nkb * kKBlock;
    LabelKernelM lk;
    RefineKernelM rk;
    if (fp16)
        full ? pick_mfma<true, true>(l.nkb, lk, rk) : pick_mfma<true, false>(l.nkb, lk, rk);
    else
        full ? pick_mfma<false, true>(l.nkb, lk, rk) : pick_mfma<false, false>(l.nkb, lk, rk);
    const size_t lds = l.table_bytes + 2 * l.norm_bytes;  // table, norms, norms + K_L
    const size_t lds_r = l.refine_lds;
    CLK_HIP_CHECK(allow_lds((const void *)lk, lds));
    CLK_HIP_CHECK(allow_lds((const void *)rk, lds_r));
    const int fast_block = 64 * fast_waves((int)l.nkb);
    const uint64_t grid = auto_grid(ctx, (const void *)lk, fast_block, lds, div_up(n, (uint64_t)64) * 64);
    {
        KernelTimer t(ctx, fp16 ? "LabelingMfmaF16" : "LabelingMfmaBF16", stream);
        // the pinned copy when it was built from these points for this strategy
        const bool pinned = ctx->mfma_pin_src == (const void *)points && ctx->mfma_pin_n == n &&
                            ctx->mfma_pin_d == d && ctx->mfma_pin_fp16 == fp16;
        const u16x8 *pin_ops = pinned ? static_cast<const u16x8 *>(ctx->mfma_pin.ptr) : nullptr;
        const f32v4 *pin_x2 =
            pinned ? reinterpret_cast<const f32v4 *>(pin_ops + pin_ops_count(n, l.nkb)) : nullptr;
        hipLaunchKernelGGL(lk, dim3((unsigned)grid), dim3(fast_block), lds, stream, points, n,
                           (uint32_t)d, (uint32_t)k, table, norm, stats, l.ntiles, prm, labels,
                           qidx, qthr, qlen, hint, pin_ops, pin_x2);
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

extern "C" int clk_labeling_pin_points(clk_context *ctx, int32_t strategy, uint64_t d, uint64_t n,
                                       const float *points, void *stream)
{
    using namespace clk;
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(strategy == CLK_LABELING_MFMA_BF16 || strategy == CLK_LABELING_MFMA_FP16,
                CLK_ERR_INVALID_ARGUMENT, "clk_labeling_pin_points: strategy must be mfma_bf16 or mfma_fp16");
    CLK_REQUIRE(d >= 1 && d <= (uint64_t)kMaxKb * kKBlock && n >= 1 && n < (1ull << 32), CLK_ERR_SIZE,
                "clk_labeling_pin_points: needs 1 <= d <= 128 and 1 <= n < 2^32");
    CLK_REQUIRE(points != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_labeling_pin_points: null points");
    const bool fp16 = strategy == CLK_LABELING_MFMA_FP16;
    const uint32_t nkb = (uint32_t)div_up(d, (uint64_t)kKBlock);
    ctx->mfma_pin_src = nullptr;  // invalid until rebuilt
    CLK_TRY(ensure_scratch(ctx->mfma_pin, pin_bytes(n, nkb)));
    u16x8 *ops = static_cast<u16x8 *>(ctx->mfma_pin.ptr);
    float *x2 = reinterpret_cast<float *>(ops + pin_ops_count(n, nkb));
    const uint64_t steps = div_up(n, (uint64_t)64);
    const dim3 grid((unsigned)div_up(steps, (uint64_t)4)), block(256);
    const hipStream_t st = as_stream(stream);
    using PinKernel = void (*)(const float *, uint64_t, uint32_t, u16x8 *, float *);
    PinKernel kern;
    switch (nkb) {
    case 1: kern = fp16 ? mfma_pin_kernel<true, 1> : mfma_pin_kernel<false, 1>; break;
    case 2: kern = fp16 ? mfma_pin_kernel<true, 2> : mfma_pin_kernel<false, 2>; break;
    case 3: kern = fp16 ? mfma_pin_kernel<true, 3> : mfma_pin_kernel<false, 3>; break;
    default: kern = fp16 ? mfma_pin_kernel<true, 4> : mfma_pin_kernel<false, 4>; break;
    }
    {
        KernelTimer t(ctx, "MfmaPinPoints", st);
        hipLaunchKernelGGL(kern, grid, block, 0, st, points, n, (uint32_t)d, ops, x2);
        CLK_HIP_CHECK(hipGetLastError());
    }
    ctx->mfma_pin_src = points;
    ctx->mfma_pin_d = d;
    ctx->mfma_pin_n = n;
    ctx->mfma_pin_fp16 = fp16;
    return CLK_OK;
}

extern "C" int clk_labeling_unpin_points(clk_context *ctx)
{
    CLK_REQUIRE(ctx != nullptr, CLK_ERR_INVALID_ARGUMENT, "null clk_context");
    ctx->mfma_pin_src = nullptr;
    return CLK_OK;
}

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
