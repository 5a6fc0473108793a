// Many independent Lloyd jobs per launch (the batch API, clkmeans_batch.h).
//
// Device state: a job table (BJob, one per slot), a tile list (each job's
// points cut into fixed tiles of tile_pts points, contiguous per job) and a
// control block.  One iteration of every active job is ONE persistent kernel
// launch: work-groups take tiles from a global counter; a tile is labeled
// with the reference arithmetic (centroids staged in LDS, sequential fma
// distance, strict '<', lowest index) and accumulated into per-wave LDS
// tables; the tile's sums and masses go to the job's FP64 partial row for
// that tile.  The work-group that completes a job's last tile (agent-scope
// release / acquire around a per-job counter) merges the job's rows in tile
// order in FP64, divides (one rounding to float, NaN for an empty cluster),
// computes the FP64 max centroid shift and advances the job.  The last
// work-group to leave resets the work counter, so the launch can be replayed
// from a HIP graph without any other node.
//
// With an MFMA labeling strategy each tile is labeled like the dense-distance
// kernels of clk_mfma.hip (the job's 16-bit centroid table staged in LDS,
// scores on the matrix cores, the margin test; undecided points of a wave
// are settled in the wave by exact FP32 distances to their candidate
// centroids), and accumulated straight from the MFMA lane layout (the 4 lanes
// of a point column own 8 features each).  The finalizing work-group also
// builds the job's table for the next iteration.
//
// Accumulation without float atomics: each wave owns an LDS table T[k][DMAX]
// (FP32) and a mass row M[k]; a lane whose point has label L writes its id to
// the work-group's owner array O[L] and reads it back; the lanes that find
// their own id -- one per distinct label -- add their point to their wave's
// T[L] / M[L] with plain read-modify-writes, the others retry (a lane of
// another wave may also win O[L]; that only costs a round).  A job with few
// clusters gets R replicas of each wave's table (R = 2..16, lanes spread over
// them), so a wave's 64 points collide less; every replica is summed in the
// tile's partial row in a fixed order.
#include "clk_common.hpp"
#include "clk_kernels.hpp"
#include "clk_mfma_device.hpp"
#include "clkmeans_batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace clk {
namespace {

struct BJob {
    const float *pts;   // [d][n]
    float *cen[2];      // [d][k], current = cen[iter & 1]
    uint32_t *labels;   // [n]
    uint32_t *masses;   // [k]
    double *part;       // [ntiles][k*(d+1)]: sums [d][k] then masses [k]
    uint64_t n;
    uint32_t d, k;
    uint32_t tile0, ntiles, tile_pts;
    uint32_t max_iter, iter;
    uint32_t done;      // tiles finished in the current iteration
    uint32_t state;     // CLK_JOB_*
    uint32_t dmax;      // register bucket of d
    double tol, shift;
    // MFMA labeling: 16-bit table [nt16][nkb][64], norms [nt16*16], stats[16]
    mfma::u16x8 *table;
    float *norm;
    float *stats;
    mfma::MfmaParams prm;
    uint32_t nt16, nkb;
    uint32_t rep;       // table replicas per wave (power of two)
};

struct BTile {
    uint32_t job, index;
};

struct BCtl {
    uint32_t next;    // work counter
    uint32_t exited;  // work-groups that left this launch
    uint32_t ntiles;  // tiles in the list
    uint32_t dry;     // 1: label and accumulate only (no merge, no job advance)
};

constexpr uint32_t kBatchMaxWaves = 4;
constexpr uint32_t kBatchPairs = 256;  // candidate pairs per wave and flush (MFMA refinement)

// LDS of the MFMA batch kernel for the largest admitted job (byte offsets):
// table | norms | W tables [kmax][pmax] | W mass rows [kmax] | owners [kmax] |
// W pair lists | W running minima [64] | W pair counters
struct BLayout {
    uint32_t kmax, pmax, waves;
    uint32_t o_norm, o_tabs, o_mass, o_owner, o_pairs, o_min, o_count, bytes;
};

inline BLayout mfma_layout_for(uint32_t waves, uint32_t kmax, uint32_t dmax_feat)
{
    BLayout L{};
    const uint32_t nt16 = (kmax + mfma::kTile - 1) / mfma::kTile;
    const uint32_t nkb = (dmax_feat + mfma::kKBlock - 1) / mfma::kKBlock;
    L.kmax = kmax;
    L.pmax = nkb * mfma::kKBlock;
    L.waves = waves;
    size_t o = (size_t)nt16 * nkb * 64 * 16;
    L.o_norm = (uint32_t)o;
    o += (size_t)nt16 * mfma::kTile * 4;
    L.o_tabs = (uint32_t)o;
    o += (size_t)waves * kmax * L.pmax * 4;
    L.o_mass = (uint32_t)o;
    o += (size_t)waves * kmax * 4;
    L.o_owner = (uint32_t)o;
    o += (size_t)kmax * 4;
    L.o_pairs = (uint32_t)o;
    o += (size_t)waves * kBatchPairs * 4;
    o = (o + 15) / 16 * 16;
    L.o_min = (uint32_t)o;
    o += (size_t)waves * 64 * 8;
    L.o_count = (uint32_t)o;
    o += (size_t)waves * 4;
    L.bytes = (uint32_t)((o + 15) / 16 * 16);
    return L;
}

// LDS: centroids [k][DMAX] | W tables [k][DMAX] | W mass rows [k] | owners [k]
__host__ __device__ inline size_t batch_lds_bytes(uint32_t waves, uint32_t k, uint32_t dm)
{
    return (size_t)k * dm * 4 * (1 + waves) + (size_t)k * 4 * (waves + 1);
}

template <int DMAX>
__device__ void batch_tile(const BJob &J, uint32_t tindex, float *lds, uint32_t waves)
{
    const uint32_t tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const uint32_t d = J.d, k = J.k;
    const uint32_t R = J.rep, rk = R * k;
    float *cen_s = lds;
    float *tabs = cen_s + (size_t)k * DMAX;
    uint32_t *mass = reinterpret_cast<uint32_t *>(tabs + (size_t)waves * rk * DMAX);
    uint32_t *owner = mass + (size_t)waves * rk;
    // this lane's replica of its wave's table: rows r*k + c
    const uint32_t roff = (lane & (R - 1)) * k;
    float *T = tabs + ((size_t)wave * rk + roff) * DMAX;
    uint32_t *M = mass + (size_t)wave * rk + roff;
    uint32_t *O = owner + roff;

    stage_centroids<DMAX>(J.cen[J.iter & 1], d, k, cen_s);
    for (uint32_t i = tid; i < waves * rk * DMAX; i += blockDim.x)
        tabs[i] = 0.0f;
    for (uint32_t i = tid; i < waves * rk; i += blockDim.x)
        mass[i] = 0u;
    __syncthreads();

    const uint64_t first = (uint64_t)tindex * J.tile_pts;
    const uint64_t end = std::min<uint64_t>(first + J.tile_pts, J.n);
    const uint64_t n = J.n;
    for (uint64_t base = first; base < end; base += blockDim.x) {
        const uint64_t p = base + tid;
        const bool valid = p < end;
        const uint64_t pc = valid ? p : first;
        float x[DMAX][1];
#pragma unroll
        for (int f = 0; f < DMAX; ++f)
            x[f][0] = (f < (int)d) ? J.pts[(uint64_t)f * n + pc] : 0.0f;
        uint32_t lab[1];
        argmin_lds<DMAX, 1>(x, cen_s, k, lab);
        if (valid)
            J.labels[p] = lab[0];
        const uint32_t L = lab[0];
        bool pending = valid;
        while (__ballot(pending) != 0) {
            if (pending)
                O[L] = tid;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const bool win = pending && O[L] == tid;
            if (win) {
                float *row = T + (size_t)L * DMAX;
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (f < (int)d)
                        row[f] = row[f] + x[f][0];
                M[L] = M[L] + 1u;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            pending = pending && !win;
        }
        (void)lane;
    }
    __syncthreads();
    // partial row of this tile (FP64, waves summed in a fixed order)
    double *row = J.part + (size_t)tindex * ((size_t)k * (d + 1));
    for (uint32_t i = tid; i < k * d; i += blockDim.x) {
        const uint32_t f = i / k, c = i % k;
        double s = 0.0;
        for (uint32_t w = 0; w < waves * R; ++w)
            s += (double)tabs[((size_t)w * k + c) * DMAX + f];
        row[i] = s;
    }
    for (uint32_t c = tid; c < k; c += blockDim.x) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < waves * R; ++w)
            m += mass[(size_t)w * k + c];
        row[(size_t)k * d + c] = (double)m;
    }
}

// Merge a job's tile rows in tile order, divide, shift, advance the job.
__device__ void batch_finalize(BJob &J, double *red)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t d = J.d, k = J.k, nt = J.ntiles;
    const size_t rl = (size_t)k * (d + 1);
    const float *old = J.cen[J.iter & 1];
    float *nxt = J.cen[(J.iter & 1) ^ 1];
    double local = 0.0;
    for (uint32_t c = tid; c < k; c += blockDim.x) {
        double m = 0.0;
        for (uint32_t t = 0; t < nt; ++t)
            m += J.part[t * rl + (size_t)k * d + c];
        double sq = 0.0;
        for (uint32_t f = 0; f < d; ++f) {
            double s = 0.0;
            for (uint32_t t = 0; t < nt; ++t)
                s += J.part[t * rl + (size_t)f * k + c];
            const float v = (float)(s / m);
            nxt[(size_t)f * k + c] = v;
            const double diff = (double)v - (double)old[(size_t)f * k + c];
            sq += diff * diff;
        }
        J.masses[c] = (uint32_t)m;
        if (m != 0.0)
            local = fmax(local, sqrt(sq));
    }
    red[tid] = local;
    __syncthreads();
    // max over blockDim.x entries (64 x waves, not always a power of two)
    uint32_t p2 = 1;
    while (p2 * 2 <= blockDim.x)
        p2 *= 2;
    if (tid < blockDim.x - p2)
        red[tid] = fmax(red[tid], red[tid + p2]);
    __syncthreads();
    for (uint32_t s = p2 / 2; s > 0; s /= 2) {
        if (tid < s)
            red[tid] = fmax(red[tid], red[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        J.shift = red[0];
        J.iter += 1;
        J.done = 0;
        if (J.iter >= J.max_iter || (J.tol > 0.0 && red[0] < J.tol))
            J.state = CLK_JOB_DONE;
    }
}

__global__ void __launch_bounds__(64 * kBatchMaxWaves)
batch_step_kernel(BJob *__restrict__ jobs, const BTile *__restrict__ tiles, BCtl *__restrict__ ctl)
{
    extern __shared__ __align__(16) float lds[];
    __shared__ uint32_t s_item;
    __shared__ int s_last;
    __shared__ double s_red[64 * kBatchMaxWaves];
    const uint32_t tid = threadIdx.x;
    const uint32_t waves = blockDim.x / 64;
    const uint32_t ntiles = ctl->ntiles;
    const bool dry = ctl->dry != 0;
    for (;;) {
        __syncthreads();
        if (tid == 0)
            s_item = atomicAdd(&ctl->next, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= ntiles)
            break;
        const BTile t = tiles[item];
        BJob &J = jobs[t.job];
        if (J.state != CLK_JOB_ACTIVE)
            continue;
        switch (J.dmax) {
        case 4: batch_tile<4>(J, t.index, lds, waves); break;
        case 8: batch_tile<8>(J, t.index, lds, waves); break;
        case 16: batch_tile<16>(J, t.index, lds, waves); break;
        case 32: batch_tile<32>(J, t.index, lds, waves); break;
        default: batch_tile<64>(J, t.index, lds, waves); break;
        }
        // publish this tile's row; the job's last tile merges
        __threadfence();
        __syncthreads();
        if (tid == 0)
            s_last = atomicAdd(&J.done, 1u) == J.ntiles - 1;
        __syncthreads();
        if (s_last) {
            __threadfence();
            if (dry) {
                if (tid == 0)
                    J.done = 0;
            } else {
                batch_finalize(J, s_red);
            }
        }
    }
    // the last work-group out resets the counters for the next launch
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(&ctl->exited, 1u) == gridDim.x - 1) {
            ctl->next = 0;
            ctl->exited = 0;
        }
    }
}

template <bool FP16, int NKB>
__device__ void batch_tile_mfma(const BJob &J, uint32_t tindex, char *lds, const BLayout &Lo)
{
    using namespace mfma;
    constexpr uint32_t P = NKB * kKBlock;
    const uint32_t tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const uint32_t q = lane % 16, h = lane / 16;
    const uint32_t waves = Lo.waves;
    const uint32_t d = J.d, k = J.k, nt16 = J.nt16;
    const uint64_t n = J.n;
    u16x8 *ltab = reinterpret_cast<u16x8 *>(lds);
    float *lnorm = reinterpret_cast<float *>(lds + Lo.o_norm);
    const uint32_t R = J.rep, rk = R * k;
    // this column's replica of its wave's table: rows r*k + c
    const uint32_t roff = (q & (R - 1)) * k;
    float *T = reinterpret_cast<float *>(lds + Lo.o_tabs) + (size_t)wave * Lo.kmax * Lo.pmax + (size_t)roff * P;
    uint32_t *M = reinterpret_cast<uint32_t *>(lds + Lo.o_mass) + (size_t)wave * Lo.kmax + roff;
    uint32_t *owner = reinterpret_cast<uint32_t *>(lds + Lo.o_owner) + roff;
    uint32_t *pl = reinterpret_cast<uint32_t *>(lds + Lo.o_pairs) + (size_t)wave * kBatchPairs;
    unsigned long long *wmin = reinterpret_cast<unsigned long long *>(lds + Lo.o_min) + (size_t)wave * 64;
    uint32_t *pc = reinterpret_cast<uint32_t *>(lds + Lo.o_count) + wave;
    const float *cen = J.cen[J.iter & 1];

    for (uint32_t e = tid; e < nt16 * NKB * 64; e += blockDim.x)
        ltab[e] = J.table[e];
    for (uint32_t c = tid; c < nt16 * kTile; c += blockDim.x)
        lnorm[c] = J.norm[c];
    for (uint32_t w = 0; w < waves; ++w) {
        float *Tw = reinterpret_cast<float *>(lds + Lo.o_tabs) + (size_t)w * Lo.kmax * Lo.pmax;
        uint32_t *Mw = reinterpret_cast<uint32_t *>(lds + Lo.o_mass) + (size_t)w * Lo.kmax;
        for (uint32_t i = tid; i < rk * P; i += blockDim.x)
            Tw[i] = 0.0f;
        for (uint32_t i = tid; i < rk; i += blockDim.x)
            Mw[i] = 0u;
    }
    __syncthreads();
    const bool table_ok = J.stats[2] == 0.0f;
    const uint32_t kmask = (1u << J.prm.key_bits) - 1u;
    const uint64_t first = (uint64_t)tindex * J.tile_pts;
    const uint64_t end = std::min<uint64_t>(first + J.tile_pts, n);
    const uint64_t nb = (end - first + 63) / 64;
    for (uint64_t bt = wave; bt < nb; bt += waves) {
        const uint64_t base = first + bt * 64;
        uint64_t pid[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const uint64_t p = base + g * 16 + q;
            pid[g] = p < end ? p : ~0ull;
        }
        PointOps<NKB> x;
        float X2[kGroups], RX2[kGroups], raw[kGroups][NKB * 8];
        load_points_keep<FP16, NKB, false>(x, X2, RX2, J.pts, n, d, pid, q, h, raw);
        const WaveBounds w = wave_bounds<NKB>(X2, RX2, J.stats, J.prm);
        uint32_t best[kGroups], second[kGroups];
        rank_all<FP16, NKB>(ltab, lnorm, nt16, lane, h, w.K, kmask, x, best, second);
        Verdict v[kGroups];
        uint32_t lab[kGroups];
        bool need = false;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            v[g] = decide(best[g], second[g], w, g, k, kmask, table_ok, J.prm);
            lab[g] = v[g].label;
            need = need || (pid[g] < n && !v[g].unique);
        }
        if (__ballot(need) != 0) {
            // exact FP32 distances to the candidates of the undecided points
            wmin[lane] = ~0ull;
            if (lane == 0)
                *pc = 0;
            wave_lds_order();
            auto flush = [&](uint32_t count) {
                for (uint32_t r = lane; r < count; r += 64) {
                    const uint32_t pr = pl[r];
                    const uint32_t slot = pr >> 20, c = pr & 0xFFFFFu;
                    const float e = exact_dist_global(J.pts, n, base + slot, cen, d, k, c);
                    if (e < FLT_MAX)  // the reference's strict '<' from FLT_MAX
                        atomicMin(&wmin[slot], ((unsigned long long)__float_as_uint(e) << 32) | c);
                }
                wave_lds_order();
                if (lane == 0)
                    *pc = 0;
                wave_lds_order();
            };
            for (uint32_t t = 0; t < nt16; ++t) {
                f32v4 acc[kGroups];
                score_tile<FP16, NKB>(acc, load_ops<NKB>(ltab, lnorm, t, lane, h), w.K, x);
#pragma unroll
                for (int g = 0; g < kGroups; ++g) {
                    uint32_t cmask = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t c = t * kTile + 4 * h + i;
                        const float sc = __uint_as_float(__float_as_uint(acc[g][i]) & ~kmask);
                        const bool cand = pid[g] < n && !v[g].unique && c < k &&
                                          (v[g].full || !(sc > v[g].thr));
                        cmask |= (uint32_t)cand << i;
                    }
                    const uint32_t nl = (uint32_t)__builtin_popcountll(__ballot(cmask != 0));
                    if (nl == 0)
                        continue;
                    const uint32_t before = *pc;
                    wave_lds_order();
                    if (before + 4 * nl > kBatchPairs)  // at most 4 pairs per such lane
                        flush(min(before, kBatchPairs));
                    while (cmask != 0) {
                        const uint32_t i = (uint32_t)__builtin_ctz(cmask);
                        cmask &= cmask - 1;
                        const uint32_t at = atomicAdd(pc, 1u);
                        pl[at] = ((uint32_t)(g * 16 + q) << 20) | (t * kTile + 4 * h + i);
                    }
                    wave_lds_order();
                }
            }
            flush(min(*pc, kBatchPairs));
#pragma unroll
            for (int g = 0; g < kGroups; ++g)
                if (pid[g] < n && !v[g].unique) {
                    const unsigned long long mn = wmin[g * 16 + q];
                    lab[g] = (mn == ~0ull) ? 0u : (uint32_t)mn;
                }
            wave_lds_order();
        }
        // labels: lane (q, h) stores group h's point
        uint32_t mine = lab[0];
        uint64_t mp = pid[0];
#pragma unroll
        for (int g = 1; g < kGroups; ++g)
            if ((uint32_t)g == h) {
                mine = lab[g];
                mp = pid[g];
            }
        if (mp < n)
            J.labels[mp] = mine;
        // accumulate: the 4 lanes of a column add their 8-feature slices
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const uint32_t L = lab[g];
            const uint32_t id = wave * 64 + g * 16 + q;
            bool pending = pid[g] < n;
            while (__ballot(pending) != 0) {
                if (pending)
                    owner[L] = id;
                wave_lds_order();
                const bool win = pending && owner[L] == id;
                if (win) {
                    float *row = T + (size_t)L * P + 8 * h;
#pragma unroll
                    for (int kb = 0; kb < NKB; ++kb) {
                        f32v4 *r4 = reinterpret_cast<f32v4 *>(row + kb * kKBlock);
                        r4[0] = r4[0] + f32v4{raw[g][kb * 8 + 0], raw[g][kb * 8 + 1], raw[g][kb * 8 + 2],
                                              raw[g][kb * 8 + 3]};
                        r4[1] = r4[1] + f32v4{raw[g][kb * 8 + 4], raw[g][kb * 8 + 5], raw[g][kb * 8 + 6],
                                              raw[g][kb * 8 + 7]};
                    }
                    if (h == 0)
                        M[L] = M[L] + 1u;
                }
                wave_lds_order();
                pending = pending && !win;
            }
        }
    }
    __syncthreads();
    double *prow = J.part + (size_t)tindex * ((size_t)k * (d + 1));
    for (uint32_t i = tid; i < k * d; i += blockDim.x) {
        const uint32_t f = i / k, c = i % k;
        double sum = 0.0;
        for (uint32_t w = 0; w < waves; ++w) {
            const float *Tw = reinterpret_cast<float *>(lds + Lo.o_tabs) + (size_t)w * Lo.kmax * Lo.pmax;
            for (uint32_t r = 0; r < R; ++r)
                sum += (double)Tw[((size_t)r * k + c) * P + f];
        }
        prow[i] = sum;
    }
    for (uint32_t c = tid; c < k; c += blockDim.x) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < waves; ++w) {
            const uint32_t *Mw = reinterpret_cast<uint32_t *>(lds + Lo.o_mass) + (size_t)w * Lo.kmax;
            for (uint32_t r = 0; r < R; ++r)
                m += Mw[(size_t)r * k + c];
        }
        prow[(size_t)k * d + c] = (double)m;
    }
}

template <bool FP16>
__global__ void __launch_bounds__(64 * kBatchMaxWaves)
batch_step_mfma_kernel(BJob *__restrict__ jobs, const BTile *__restrict__ tiles, BCtl *__restrict__ ctl,
                       BLayout Lo)
{
    extern __shared__ __align__(16) char lds_b[];
    __shared__ uint32_t s_item;
    __shared__ int s_last;
    __shared__ double s_red[64 * kBatchMaxWaves];
    const uint32_t tid = threadIdx.x;
    const uint32_t ntiles = ctl->ntiles;
    const bool dry = ctl->dry != 0;
    for (;;) {
        __syncthreads();
        if (tid == 0)
            s_item = atomicAdd(&ctl->next, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= ntiles)
            break;
        const BTile t = tiles[item];
        BJob &J = jobs[t.job];
        if (J.state != CLK_JOB_ACTIVE)
            continue;
        if (J.nkb == 1)
            batch_tile_mfma<FP16, 1>(J, t.index, lds_b, Lo);
        else
            batch_tile_mfma<FP16, 2>(J, t.index, lds_b, Lo);
        __threadfence();
        __syncthreads();
        if (tid == 0)
            s_last = atomicAdd(&J.done, 1u) == J.ntiles - 1;
        __syncthreads();
        if (s_last) {
            __threadfence();
            if (dry) {
                if (tid == 0)
                    J.done = 0;
            } else {
                const uint32_t nxt = (J.iter & 1) ^ 1;
                batch_finalize(J, s_red);
                __syncthreads();
                mfma::build_table_block<FP16>(J.cen[nxt], J.d, J.k, J.nt16, J.nkb, J.table, J.norm,
                                              J.stats, s_red);
            }
        }
    }
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(&ctl->exited, 1u) == gridDim.x - 1) {
            ctl->next = 0;
            ctl->exited = 0;
        }
    }
}

// The table of a newly admitted job (from its initial centroids).
template <bool FP16>
__global__ void __launch_bounds__(256) batch_build_table_kernel(BJob *__restrict__ jobs, uint32_t slot)
{
    __shared__ double red[256];
    BJob &J = jobs[slot];
    mfma::build_table_block<FP16>(J.cen[0], J.d, J.k, J.nt16, J.nkb, J.table, J.norm, J.stats, red);
}

uint32_t bucket_of(uint32_t d)
{
    uint32_t b = 4;
    while (b < d)
        b *= 2;
    return b;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace
}  // namespace clk

using clk::BCtl;
using clk::BJob;
using clk::BTile;

struct clk_batch {
    clk_context *ctx = nullptr;
    clk_batch_config cfg{};
    uint32_t waves = 1;
    size_t lds = 0;
    bool mfma = false, fp16 = false;
    clk::BLayout layout{};
    BJob *d_jobs = nullptr;
    BCtl *d_ctl = nullptr;
    BTile *d_tiles = nullptr;
    uint32_t tile_cap = 0;
    char *arena = nullptr;
    // host mirror of the job table (device-side fields iter/state/shift are
    // read back by poll / take)
    std::vector<BJob> jobs;
    std::vector<uint64_t> off, bytes;  // arena ranges per slot
    uint32_t tile_work = 0;
    bool dirty = true;  // tile list / graph must be rebuilt
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    uint32_t grid = 0;
};

namespace {

uint64_t job_bytes(uint64_t n, uint32_t d, uint32_t k, uint32_t tile_pts)
{
    const uint64_t nt = clk::div_up(std::max<uint64_t>(n, 1), tile_pts);
    const uint64_t nt16 = clk::div_up(k, clk::mfma::kTile), nkb = clk::div_up(d, clk::mfma::kKBlock);
    uint64_t b = 0;
    b += clk::align_up(nt16 * nkb * 64 * 16, 256);        // 16-bit table
    b += clk::align_up(nt16 * clk::mfma::kTile * 4, 256);  // norms
    b += 256;                                             // stats
    b += clk::align_up(n * d * 4, 256);                   // points
    b += clk::align_up(n * 4, 256);                       // labels
    b += 2 * clk::align_up((uint64_t)d * k * 4, 256);     // centroids x2
    b += clk::align_up((uint64_t)k * 4, 256);             // masses
    b += clk::align_up(nt * (uint64_t)k * (d + 1) * 8, 256);  // partial rows
    return b;
}

// Points per tile: tile_work points per cluster (so the FP64 partial rows,
// k*(d+1) values per tile, stay a small fraction of the point data), between
// 1024 and 16384.
uint32_t tile_points(uint32_t tile_work, uint32_t d, uint32_t k)
{
    (void)d;
    uint64_t t = (uint64_t)tile_work * k;
    t = std::max<uint64_t>(1024, std::min<uint64_t>(t, 16384));
    return (uint32_t)(t / 256 * 256);
}

// First fit over the arena (slots hold [off, off+bytes)).
bool arena_alloc(clk_batch *b, uint64_t need, uint64_t *where)
{
    std::vector<std::pair<uint64_t, uint64_t>> used;
    for (size_t s = 0; s < b->jobs.size(); ++s)
        if (b->jobs[s].state != CLK_JOB_FREE)
            used.push_back({b->off[s], b->bytes[s]});
    std::sort(used.begin(), used.end());
    uint64_t pos = 0;
    for (auto &u : used) {
        if (u.first >= pos + need)
            break;
        pos = std::max(pos, u.first + u.second);
    }
    if (pos + need > b->cfg.arena_bytes)
        return false;
    *where = pos;
    return true;
}

int rebuild(clk_batch *b, hipStream_t stream)
{
    std::vector<BTile> tiles;
    for (uint32_t s = 0; s < b->jobs.size(); ++s) {
        BJob &J = b->jobs[s];
        if (J.state != CLK_JOB_ACTIVE)
            continue;
        J.tile0 = (uint32_t)tiles.size();
        for (uint32_t t = 0; t < J.ntiles; ++t)
            tiles.push_back({s, t});
    }
    if (tiles.size() > b->tile_cap) {
        if (b->d_tiles)
            CLK_HIP_CHECK(hipFree(b->d_tiles));
        b->tile_cap = (uint32_t)std::max<size_t>(tiles.size(), 1024) * 2;
        CLK_HIP_CHECK(hipMalloc(&b->d_tiles, b->tile_cap * sizeof(BTile)));
    }
    if (!tiles.empty())
        CLK_HIP_CHECK(hipMemcpyAsync(b->d_tiles, tiles.data(), tiles.size() * sizeof(BTile),
                                     hipMemcpyHostToDevice, stream));
    BCtl ctl{0, 0, (uint32_t)tiles.size(), 0};
    CLK_HIP_CHECK(hipMemcpyAsync(b->d_ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice, stream));
    CLK_HIP_CHECK(hipStreamSynchronize(stream));  // tiles / ctl are host temporaries
    const void *kern = !b->mfma ? (const void *)clk::batch_step_kernel
                     : b->fp16  ? (const void *)clk::batch_step_mfma_kernel<true>
                                : (const void *)clk::batch_step_mfma_kernel<false>;
    b->grid = tiles.empty() ? 0
                            : (uint32_t)clk::auto_grid(b->ctx, kern, (int)(64 * b->waves), b->lds,
                                                       (uint64_t)tiles.size() * 64 * b->waves);
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    b->dirty = false;
    return CLK_OK;
}

int launch_step(clk_batch *b, hipStream_t stream)
{
    if (!b->mfma)
        hipLaunchKernelGGL(clk::batch_step_kernel, dim3(b->grid), dim3(64 * b->waves), b->lds, stream,
                           b->d_jobs, b->d_tiles, b->d_ctl);
    else if (b->fp16)
        hipLaunchKernelGGL(clk::batch_step_mfma_kernel<true>, dim3(b->grid), dim3(64 * b->waves), b->lds,
                           stream, b->d_jobs, b->d_tiles, b->d_ctl, b->layout);
    else
        hipLaunchKernelGGL(clk::batch_step_mfma_kernel<false>, dim3(b->grid), dim3(64 * b->waves), b->lds,
                           stream, b->d_jobs, b->d_tiles, b->d_ctl, b->layout);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

}  // namespace

extern "C" uint64_t clk_batch_job_bytes(const clk_batch_config *cfg, uint64_t n, uint32_t d, uint32_t k)
{
    const uint32_t tw = (cfg && cfg->tile_work) ? cfg->tile_work : 64;
    return job_bytes(n, d, k, tile_points(tw, d, k));
}

extern "C" int clk_batch_create(clk_context *ctx, const clk_batch_config *cfg, clk_batch **out)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(cfg && out, CLK_ERR_INVALID_ARGUMENT, "clk_batch_create: null argument");
    CLK_REQUIRE(cfg->max_jobs >= 1 && cfg->arena_bytes > 0, CLK_ERR_SIZE, "clk_batch_create: empty batch");
    CLK_REQUIRE(cfg->max_d >= 1 && cfg->max_d <= 64 && cfg->max_k >= 1, CLK_ERR_SIZE,
                "clk_batch_create: max_d must be 1..64, max_k >= 1");
    *out = nullptr;
    clk_batch *b = new clk_batch;
    b->ctx = ctx;
    b->cfg = *cfg;
    b->tile_work = cfg->tile_work ? cfg->tile_work : 64;
    CLK_REQUIRE(cfg->labeling == CLK_LABELING_UNROLL_VECTOR || cfg->labeling == CLK_LABELING_MFMA_BF16 ||
                    cfg->labeling == CLK_LABELING_MFMA_FP16,
                CLK_ERR_INVALID_ARGUMENT, "clk_batch_create: unknown labeling strategy id");
    b->mfma = cfg->labeling != CLK_LABELING_UNROLL_VECTOR;
    b->fp16 = cfg->labeling == CLK_LABELING_MFMA_FP16;
    const uint32_t dm = clk::bucket_of(cfg->max_d);
    b->waves = 0;
    for (uint32_t w = clk::kBatchMaxWaves; w >= 1; --w) {
        const size_t need = b->mfma ? clk::mfma_layout_for(w, cfg->max_k, cfg->max_d).bytes
                                    : clk::batch_lds_bytes(w, cfg->max_k, dm);
        if (need + 4096 <= ctx->lds_per_block) {  // + the kernel's static LDS
            b->waves = w;
            break;
        }
    }
    if (b->waves == 0) {
        delete b;
        clk::set_error("clk_batch_create: max_k x max_d too large for the work-group's LDS");
        return CLK_ERR_UNSUPPORTED;
    }
    if (b->mfma) {
        b->layout = clk::mfma_layout_for(b->waves, cfg->max_k, cfg->max_d);
        b->lds = b->layout.bytes;
        const int rc = clk::mfma_subnormal_flags(ctx, nullptr);
        if (rc != CLK_OK) {
            delete b;
            return rc;
        }
    } else {
        b->lds = clk::batch_lds_bytes(b->waves, cfg->max_k, dm);
    }
    if (b->lds > 65536) {
        const void *kern = !b->mfma ? (const void *)clk::batch_step_kernel
                         : b->fp16  ? (const void *)clk::batch_step_mfma_kernel<true>
                                    : (const void *)clk::batch_step_mfma_kernel<false>;
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds);
        if (e != hipSuccess) {
            delete b;
            clk::set_error(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
            return CLK_ERR_HIP;
        }
    }
    b->jobs.assign(cfg->max_jobs, BJob{});
    b->off.assign(cfg->max_jobs, 0);
    b->bytes.assign(cfg->max_jobs, 0);
    hipError_t e1 = hipMalloc(&b->d_jobs, cfg->max_jobs * sizeof(BJob));
    hipError_t e2 = hipMalloc(&b->d_ctl, sizeof(BCtl));
    hipError_t e3 = hipMalloc(&b->arena, cfg->arena_bytes);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        clk_batch_destroy(b);
        clk::set_error("clk_batch_create: device allocation failed");
        return CLK_ERR_NOMEM;
    }
    CLK_HIP_CHECK(hipMemset(b->d_jobs, 0, cfg->max_jobs * sizeof(BJob)));
    BCtl ctl{0, 0, 0, 0};
    CLK_HIP_CHECK(hipMemcpy(b->d_ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice));
    *out = b;
    return CLK_OK;
}

extern "C" int clk_batch_destroy(clk_batch *b)
{
    if (!b)
        return CLK_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    if (b->d_jobs)
        (void)hipFree(b->d_jobs);
    if (b->d_ctl)
        (void)hipFree(b->d_ctl);
    if (b->d_tiles)
        (void)hipFree(b->d_tiles);
    if (b->arena)
        (void)hipFree(b->arena);
    delete b;
    return CLK_OK;
}

extern "C" int clk_batch_add(clk_batch *b, uint64_t n, uint32_t d, uint32_t k, const float *host_points,
                             const float *host_centroids, uint32_t max_iterations, double tolerance,
                             int32_t *slot, void *stream)
{
    CLK_REQUIRE(b && slot, CLK_ERR_INVALID_ARGUMENT, "clk_batch_add: null argument");
    CLK_TRY(clk::enter(b->ctx));
    CLK_REQUIRE(n >= 1 && n < (1ull << 32), CLK_ERR_SIZE, "clk_batch_add: n must be 1..2^32-1");
    CLK_REQUIRE(d >= 1 && d <= b->cfg.max_d && k >= 1 && k <= b->cfg.max_k, CLK_ERR_SIZE,
                "clk_batch_add: d or k beyond the batch's max_d / max_k");
    CLK_REQUIRE(host_points && host_centroids, CLK_ERR_INVALID_ARGUMENT, "clk_batch_add: null data");
    *slot = -1;
    int32_t s = -1;
    for (uint32_t i = 0; i < b->jobs.size(); ++i)
        if (b->jobs[i].state == CLK_JOB_FREE) {
            s = (int32_t)i;
            break;
        }
    if (s < 0) {
        clk::set_error("clk_batch_add: no free slot");
        return CLK_ERR_NOMEM;
    }
    const uint32_t tp = tile_points(b->tile_work, d, k);
    const uint64_t need = job_bytes(n, d, k, tp);
    uint64_t at = 0;
    if (!arena_alloc(b, need, &at)) {
        clk::set_error("clk_batch_add: arena full");
        return CLK_ERR_NOMEM;
    }
    hipStream_t st = clk::as_stream(stream);
    char *p = b->arena + at;
    BJob J{};
    J.nt16 = (uint32_t)clk::div_up(k, clk::mfma::kTile);
    J.nkb = (uint32_t)clk::div_up(d, clk::mfma::kKBlock);
    J.table = reinterpret_cast<clk::mfma::u16x8 *>(p);
    p += clk::align_up((uint64_t)J.nt16 * J.nkb * 64 * 16, 256);
    J.norm = reinterpret_cast<float *>(p);
    p += clk::align_up((uint64_t)J.nt16 * clk::mfma::kTile * 4, 256);
    J.stats = reinterpret_cast<float *>(p);
    p += 256;
    J.prm = clk::mfma::make_params(b->fp16, d, k, (b->ctx->mfma_subnormals & (b->fp16 ? 1 : 2)) != 0);
    {
        // replicas: R*k rows must fit the rows and cells sized for max_k x max_d
        const uint64_t p_job = b->mfma ? (uint64_t)J.nkb * clk::mfma::kKBlock : clk::bucket_of(d);
        const uint64_t p_max = b->mfma ? (uint64_t)b->layout.pmax : clk::bucket_of(b->cfg.max_d);
        const uint64_t lim = std::min<uint64_t>(std::min<uint64_t>(16, b->cfg.max_k / k),
                                                (uint64_t)b->cfg.max_k * p_max / ((uint64_t)k * p_job));
        J.rep = 1;
        while (J.rep * 2 <= lim)
            J.rep *= 2;
    }
    J.pts = reinterpret_cast<const float *>(p);
    p += clk::align_up(n * d * 4, 256);
    J.labels = reinterpret_cast<uint32_t *>(p);
    p += clk::align_up(n * 4, 256);
    J.cen[0] = reinterpret_cast<float *>(p);
    p += clk::align_up((uint64_t)d * k * 4, 256);
    J.cen[1] = reinterpret_cast<float *>(p);
    p += clk::align_up((uint64_t)d * k * 4, 256);
    J.masses = reinterpret_cast<uint32_t *>(p);
    p += clk::align_up((uint64_t)k * 4, 256);
    J.part = reinterpret_cast<double *>(p);
    J.n = n;
    J.d = d;
    J.k = k;
    J.tile_pts = tp;
    J.ntiles = (uint32_t)clk::div_up(n, tp);
    J.max_iter = max_iterations;
    J.iter = 0;
    J.done = 0;
    J.state = max_iterations > 0 ? CLK_JOB_ACTIVE : CLK_JOB_DONE;
    J.dmax = clk::bucket_of(d);
    J.tol = tolerance;
    J.shift = 0.0;
    CLK_HIP_CHECK(hipMemcpyAsync(const_cast<float *>(J.pts), host_points, n * d * 4, hipMemcpyHostToDevice, st));
    CLK_HIP_CHECK(hipMemcpyAsync(J.cen[0], host_centroids, (size_t)d * k * 4, hipMemcpyHostToDevice, st));
    CLK_HIP_CHECK(hipMemsetAsync(J.masses, 0, (size_t)k * 4, st));
    CLK_HIP_CHECK(hipMemsetAsync(J.labels, 0, n * 4, st));
    CLK_HIP_CHECK(hipMemcpyAsync(b->d_jobs + s, &J, sizeof(BJob), hipMemcpyHostToDevice, st));
    if (b->mfma) {
        if (b->fp16)
            hipLaunchKernelGGL(clk::batch_build_table_kernel<true>, dim3(1), dim3(256), 0, st, b->d_jobs,
                               (uint32_t)s);
        else
            hipLaunchKernelGGL(clk::batch_build_table_kernel<false>, dim3(1), dim3(256), 0, st, b->d_jobs,
                               (uint32_t)s);
        CLK_HIP_CHECK(hipGetLastError());
    }
    CLK_HIP_CHECK(hipStreamSynchronize(st));  // host_points may be reused at once
    b->jobs[s] = J;
    b->off[s] = at;
    b->bytes[s] = need;
    b->dirty = true;
    *slot = s;
    return CLK_OK;
}

extern "C" int clk_batch_step(clk_batch *b, void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_step: null batch");
    CLK_TRY(clk::enter(b->ctx));
    hipStream_t st = clk::as_stream(stream);
    if (b->dirty)
        CLK_TRY(rebuild(b, st));
    if (b->grid == 0)
        return CLK_OK;
    if (!b->exec || b->graph_stream != st) {
        if (b->exec) {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
        if (b->graph) {
            (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
        }
        if (st == nullptr) {  // capture needs a real stream: launch directly
            clk::KernelTimer t(b->ctx, "BatchStep", st);
            return launch_step(b, st);
        }
        CLK_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = launch_step(b, st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        CLK_TRY(rc);
        CLK_HIP_CHECK(e);
        b->graph = g;
        CLK_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graph_stream = st;
    }
    clk::KernelTimer t(b->ctx, "BatchStep", st);
    CLK_HIP_CHECK(hipGraphLaunch(b->exec, st));
    return CLK_OK;
}

extern "C" int clk_batch_poll(clk_batch *b, uint32_t *states, uint32_t *active, void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_poll: null batch");
    CLK_TRY(clk::enter(b->ctx));
    hipStream_t st = clk::as_stream(stream);
    CLK_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<BJob> dev(b->jobs.size());
    CLK_HIP_CHECK(hipMemcpy(dev.data(), b->d_jobs, dev.size() * sizeof(BJob), hipMemcpyDeviceToHost));
    uint32_t a = 0;
    for (size_t s = 0; s < dev.size(); ++s) {
        uint32_t state = b->jobs[s].state == CLK_JOB_FREE ? CLK_JOB_FREE : dev[s].state;
        if (state == CLK_JOB_DONE && b->jobs[s].state == CLK_JOB_ACTIVE) {
            b->jobs[s].state = CLK_JOB_DONE;  // leaves the tile list
            b->dirty = true;
        }
        if (states)
            states[s] = state;
        a += state == CLK_JOB_ACTIVE;
    }
    if (active)
        *active = a;
    return CLK_OK;
}

extern "C" int clk_batch_take(clk_batch *b, int32_t slot, float *host_centroids, uint32_t *host_labels,
                              uint32_t *host_masses, uint32_t *iterations, double *shift, void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_take: null batch");
    CLK_TRY(clk::enter(b->ctx));
    CLK_REQUIRE(slot >= 0 && (uint32_t)slot < b->jobs.size() && b->jobs[slot].state != CLK_JOB_FREE,
                CLK_ERR_INVALID_ARGUMENT, "clk_batch_take: no job in this slot");
    hipStream_t st = clk::as_stream(stream);
    CLK_HIP_CHECK(hipStreamSynchronize(st));
    BJob J;
    CLK_HIP_CHECK(hipMemcpy(&J, b->d_jobs + slot, sizeof(BJob), hipMemcpyDeviceToHost));
    if (host_centroids)
        CLK_HIP_CHECK(hipMemcpy(host_centroids, J.cen[J.iter & 1], (size_t)J.d * J.k * 4, hipMemcpyDeviceToHost));
    if (host_labels)
        CLK_HIP_CHECK(hipMemcpy(host_labels, J.labels, J.n * 4, hipMemcpyDeviceToHost));
    if (host_masses)
        CLK_HIP_CHECK(hipMemcpy(host_masses, J.masses, (size_t)J.k * 4, hipMemcpyDeviceToHost));
    if (iterations)
        *iterations = J.iter;
    if (shift)
        *shift = J.shift;
    const bool was_active = b->jobs[slot].state == CLK_JOB_ACTIVE;
    J.state = CLK_JOB_FREE;
    CLK_HIP_CHECK(hipMemcpy(b->d_jobs + slot, &J, sizeof(BJob), hipMemcpyHostToDevice));
    b->jobs[slot].state = CLK_JOB_FREE;
    if (was_active)
        b->dirty = true;
    return CLK_OK;
}

extern "C" int clk_batch_job_buffers(clk_batch *b, int32_t slot, const float **points, float **centroids,
                                     uint32_t **labels)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_job_buffers: null batch");
    CLK_TRY(clk::enter(b->ctx));
    CLK_REQUIRE(slot >= 0 && (uint32_t)slot < b->jobs.size() && b->jobs[slot].state != CLK_JOB_FREE,
                CLK_ERR_INVALID_ARGUMENT, "clk_batch_job_buffers: no job in this slot");
    BJob J;
    CLK_HIP_CHECK(hipMemcpy(&J, b->d_jobs + slot, sizeof(BJob), hipMemcpyDeviceToHost));
    if (points)
        *points = J.pts;
    if (centroids)
        *centroids = J.cen[J.iter & 1];
    if (labels)
        *labels = J.labels;
    return CLK_OK;
}

extern "C" int clk_batch_run_serial(clk_batch *b, void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_run_serial: null batch");
    CLK_TRY(clk::enter(b->ctx));
    hipStream_t st = clk::as_stream(stream);
    CLK_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<BJob> dev(b->jobs.size());
    CLK_HIP_CHECK(hipMemcpy(dev.data(), b->d_jobs, dev.size() * sizeof(BJob), hipMemcpyDeviceToHost));
    clk_kernel_config cfg;
    CLK_TRY(clk_default_config(b->ctx, CLK_STAGE_FUSED, CLK_FUSED_CLUSTER_MERGE, &cfg));
    double *shift_d = nullptr;
    CLK_HIP_CHECK(hipMalloc(&shift_d, sizeof(double)));
    int rc = CLK_OK;
    for (size_t s = 0; s < dev.size() && rc == CLK_OK; ++s) {
        BJob &J = dev[s];
        if (b->jobs[s].state == CLK_JOB_FREE || J.state != CLK_JOB_ACTIVE)
            continue;
        // the job's partial rows double as its FP64 sums buffer [d][k]
        while (J.state == CLK_JOB_ACTIVE) {
            float *cur = J.cen[J.iter & 1], *nxt = J.cen[(J.iter & 1) ^ 1];
            rc = clk_lloyd_step(b->ctx, &cfg, J.d, J.n, J.k, J.pts, cur, nxt, J.part, J.labels,
                                J.masses, J.tol > 0.0 ? shift_d : nullptr, st);
            if (rc != CLK_OK)
                break;
            J.iter += 1;
            if (J.tol > 0.0) {
                hipError_t e = hipMemcpyAsync(&J.shift, shift_d, sizeof(double), hipMemcpyDeviceToHost, st);
                if (e == hipSuccess)
                    e = hipStreamSynchronize(st);
                if (e != hipSuccess) {
                    clk::set_error(std::string("clk_batch_run_serial: ") + hipGetErrorString(e));
                    rc = CLK_ERR_HIP;
                    break;
                }
            }
            if (J.iter >= J.max_iter || (J.tol > 0.0 && J.shift < J.tol))
                J.state = CLK_JOB_DONE;
        }
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(shift_d);
    CLK_TRY(rc);
    CLK_HIP_CHECK(hipMemcpy(b->d_jobs, dev.data(), dev.size() * sizeof(BJob), hipMemcpyHostToDevice));
    return CLK_OK;
}

extern "C" int clk_batch_autotune(clk_batch *b, uint32_t *chosen, void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_autotune: null batch");
    CLK_TRY(clk::enter(b->ctx));
    (void)stream;
    if (chosen)
        *chosen = b->tile_work;
    clk::set_error("clk_batch_autotune: not available yet");
    return CLK_ERR_UNSUPPORTED;
}
