// Many independent Lloyd jobs per launch (the batch API, clkmeans_batch.h).
//
// Device state: a job table (BJob, one per slot), a tile list (each job's
// points cut into fixed tiles of tile_pts points, contiguous per job), the
// list of active slots and a control block.  One iteration of every active
// job is a few launches, replayed from one HIP graph.  The tile kernel is
// persistent: work-groups take tiles from a global counter (the next one is
// requested while the current one is processed); a tile is labeled with the
// reference arithmetic (centroids staged in LDS, sequential fma distance,
// strict '<', lowest index) and accumulated into per-wave LDS tables, and
// the tile's sums and masses go to the job's FP64 partial row for that tile.
// The merge kernel (one work-group per element slice of a job's row) then
// sums the job's rows in a fixed order in FP64, and the finish kernel (one
// work-group per active job) divides (one rounding to float, NaN for an
// empty cluster), computes the FP64 max centroid shift, advances the job and
// resets the work counters.  The kernel boundaries publish the partial rows:
// no tile needs a device-scope fence, which on this GPU writes back and
// invalidates the L2 of the tile's XCD.  Jobs fall into size classes (FP32
// or MFMA labeling, d bucket, k up to 64 or above), one tile-kernel launch
// each, so that a class's LDS and registers fit its own jobs.
//
// With an MFMA labeling strategy each tile is labeled like the dense-distance
// kernels of clk_mfma.hip (the job's 16-bit centroid table staged in LDS,
// scores on the matrix cores, the margin test; undecided points of a wave
// are settled in the wave by exact FP32 distances to their candidate
// centroids).  The finish kernel also builds the job's table for the next
// iteration.
//
// Exact accumulation (clk_grid.hpp): every coordinate is rounded to its grid
// integer (a per-job, per-feature power-of-two unit).  For every batch of 64
// points a wave sorts its (label, lane) keys (a bitonic network over the
// lanes), so the points of one label form a run of lanes; an integer
// segmented scan along the sorted lanes (DPP row shifts and row broadcasts)
// leaves each run's sum on its last lane (|sum| < 64 * 2^25 fits an int32),
// which adds it to the work-group's int64 accumulator A[d][k] with one
// ds_add_u64 per feature -- one per distinct label.  Integer addition is
// exact, so a tile's row, and a job's merged sums, are the same for every
// tile size, wave count and job mix, and equal those of clk_lloyd_step.
#include "clk_common.hpp"
#include "clk_grid.hpp"
#include "clk_kernels.hpp"
#include "clk_mfma_device.hpp"
#include "clkmeans_batch.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <vector>

namespace clk {
namespace {

struct BJob {
    // status, first so that a strided copy of 16 bytes per record reads it
    uint32_t iter;      // iterations done
    uint32_t state;     // CLK_JOB_*
    double shift;       // max centroid shift of the last iteration
    const float *pts;   // [d][n]
    float *cen[2];      // [d][k], current = cen[iter & 1]
    uint32_t *labels;   // [n]
    uint32_t *masses;   // [k]
    double *part;       // [ntiles][k*(d+1)]: sums [d][k] then masses [k]
    uint64_t n;
    uint32_t d, k;
    uint32_t tile0, ntiles, tile_pts;
    uint32_t max_iter;
    uint32_t dmax;      // register bucket of d
    uint32_t kind;      // 0: FP32 labeling, 1: MFMA labeling
    double tol;
    // MFMA labeling: 16-bit table [nt16][nkb][64], norms [nt16*16], stats[16]
    mfma::u16x8 *table;
    float *norm;
    float *stats;
    mfma::MfmaParams prm;
    uint32_t nt16, nkb;
    // FP32-scan jobs: keep thresholds [k] of the current centroids
    // (keep_threshold); valid when keep_ok (set by the finish kernel, clear
    // at admission).  (The MFMA tiles do not use them: a whole wave of 64
    // points seldom passes there, and the test measured slower.)
    float *keep;
    uint32_t keep_ok;
    // FP32-scan jobs with d <= 8: lb[n], per point a lower bound on its
    // distance to every centroid but its label's (second_bound), rewritten
    // by every tile pass; valid when bound_ok (set by the finish kernel after
    // a batch iteration, clear at admission and after an iteration outside
    // the batch).  drift: an upper bound on how far any centroid that can
    // still be chosen moved in the last iteration (batch_finalize).
    uint32_t bound_ok;
    float *lb;
    float drift;
    uint32_t pad1;
    // MFMA jobs with d <= 32: the current centroids as rows [k][cen_t_w]
    // (zero-padded to 4/8/16/32 floats), refreshed by the finish kernel, for
    // the keep test and the exact rescan of the tile kernel
    float *cen_t;
    uint32_t cen_t_w;
    uint32_t tile_base;  // (host) the tile size of the batch's tile_work; rebuild may use smaller tiles
    // exact sums (clk_grid.hpp): the per-feature max |x| bits of the job's
    // points (its grid), and per cell [d][k] the flags of off-grid values
    // (set by the tiles, consumed and cleared by the finish kernel)
    uint32_t *gmax;
    uint32_t *flags;
};

struct BTile {
    uint32_t job, index;
};

struct BCtl {
    uint32_t next;    // work counter of the tile kernel
    uint32_t ntiles;  // tiles in the list
    uint32_t dry;     // 1: the finish kernel only resets the counters (no job advances)
    uint32_t pad;
};

constexpr uint32_t kBatchMaxWaves = 4;      // FP32-scan tile kernels
constexpr uint32_t kBatchMfmaWaves = 8;      // MFMA tile kernels with one K-block (two waves per SIMD)
constexpr uint32_t kBatchPairs = 256;  // candidate pairs per wave and flush (MFMA refinement)

// Global-memory (address space 1) views for pointers read from job records.
typedef __attribute__((address_space(1))) const float gfloat;
typedef __attribute__((address_space(1))) const uint32_t guint;
typedef __attribute__((address_space(1))) const double gdouble;
typedef __attribute__((address_space(1))) const mfma::f32v4 gf32v4;

// The exact accumulator of a tile in LDS (byte offsets from its start):
// A [d][k] int64 sums of grid integers | counts [k] | grid scales [smax].
struct BAcc {
    long long *A;
    uint32_t *cnt;
    float *scl;
};

__host__ __device__ inline size_t acc_bytes(uint32_t k, uint32_t kd, uint32_t smax)
{
    return (size_t)kd * 8 + (size_t)((k + 3) & ~3u) * 4 + (size_t)((smax + 3) & ~3u) * 4;
}

// Carve the accumulator of a job (d, k) at p (8-byte aligned), zero it and
// stage the job's grid scales (callers add a barrier).
__device__ inline BAcc acc_carve(char *p, const BJob &J, uint32_t smax)
{
    BAcc a;
    const uint32_t d = J.d, k = J.k;
    a.A = reinterpret_cast<long long *>(p);
    a.cnt = reinterpret_cast<uint32_t *>(a.A + (size_t)d * k);
    a.scl = reinterpret_cast<float *>(a.cnt + ((k + 3) & ~3u));
    const uint32_t bits = grid_bits(J.n);
    for (uint32_t i = threadIdx.x; i < d * k; i += blockDim.x)
        a.A[i] = 0;
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x)
        a.cnt[i] = 0u;
    for (uint32_t f = threadIdx.x; f < smax; f += blockDim.x)
        a.scl[f] = f < d ? grid_scale(J.gmax[f], bits) : 1.0f;
    return a;
}

// LDS of the MFMA batch kernel for the jobs of a class (byte offsets):
// table | norms | the exact accumulator (largest k, largest k*d) | W pair
// lists | W running minima [64] | W pair counters
struct BLayout {
    uint32_t kmax, waves;
    uint32_t o_norm, o_acc, o_pairs, o_min, o_count, bytes;
};

inline BLayout mfma_layout_for(uint32_t waves, uint32_t kmax, uint32_t kdmax, uint32_t dmax_feat)
{
    BLayout L{};
    const uint32_t nt16 = (kmax + mfma::kTile - 1) / mfma::kTile;
    const uint32_t nkb = (dmax_feat + mfma::kKBlock - 1) / mfma::kKBlock;
    L.kmax = kmax;
    L.waves = waves;
    size_t o = (size_t)nt16 * nkb * 64 * 16;
    L.o_norm = (uint32_t)o;
    o += (size_t)nt16 * mfma::kTile * 4;
    L.o_acc = (uint32_t)o;  // 16-byte aligned
    o += acc_bytes(kmax, kdmax, nkb * mfma::kKBlock);
    o = (o + 15) / 16 * 16;
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

// Floats between the staged centroid rows of the plain kernel: rows of 8, 16
// and 32 features get 4 floats of padding, so the lanes of a wave reading
// different centroids (the keep test: each lane its previous label's row;
// the wave-wide rescan: each lane another centroid) hit different LDS banks
// with their 128-bit reads, where unpadded rows start on the same few banks.
__host__ __device__ constexpr uint32_t cen_stride(uint32_t dm) { return (dm == 8 || dm == 16 || dm == 32) ? dm + 4 : dm; }

// LDS of the plain kernel: the exact accumulator (largest k, largest k*d) |
// centroids [k][cen_stride(DMAX)] | W rescan lists [64]
__host__ __device__ inline size_t batch_lds_bytes(uint32_t waves, uint32_t k, uint32_t kd, uint32_t dm)
{
    return acc_bytes(k, kd, dm) + (size_t)k * cen_stride(dm) * 4 + (size_t)waves * 64 * 4;
}

// ---- second-distance bound of the d <= 8 tiles -----------------------------
//
// A scan that returns label a also gives S, the smallest FP32 distance D(x, c)
// over the other centroids (capped at FLT_MAX; NaN distances, which never win
// the strict '<' scan, skipped).  D is the reference's fma sequence: within a
// relative (d+3)*2^-24 of the exact ||x - c||^2, plus at most d*2^-126 from
// underflow.  So with m = 2^-10, B = sqrt(S (1-m)) (0 below 1e-20) is at most
// the exact distance from x to every centroid but c_a.  When every centroid
// that can still be chosen then moves by at most `drift` (exact Euclidean,
// rounded up), L = (B - drift) (1 - 2^-20) (rounded values; 0 if negative)
// bounds the new distances the same way, and so on each iteration.  A point
// whose FP32 distance to its label's new centroid is below L^2 (1-m) keeps
// the label: every other D(x, c) >= L^2 (1 - (d+3) 2^-24) - d 2^-126 is
// larger (m dwarfs the rounding terms; thresholds below 1e-20 keep nothing),
// so c_a is the strict minimum, below FLT_MAX, and the scan returns it
// (kmeans_naive.cpp:49-63).  A centroid with a NaN or inf coordinate has
// D(x, c) NaN or >= FLT_MAX and is never chosen over such a point.
constexpr float kBoundMargin = 1.0f - 0x1p-10f;
constexpr float kBoundShrink = 1.0f - 0x1p-20f;
constexpr float kBoundMin = 1e-20f;

__device__ __forceinline__ float second_bound(float s)
{
    const float t = s * kBoundMargin;
    return t >= kBoundMin ? __builtin_sqrtf(t) : 0.0f;
}

// The running (minimum, second minimum) of the scan: strict '<' replaces the
// minimum, whose old value (or the new distance) goes into the second one;
// fminf skips a NaN distance.
__device__ __forceinline__ void scan_step2(float dist, uint32_t c, float &best, uint32_t &bi, float &sec)
{
    const bool lt = dist < best;
    sec = fminf(sec, lt ? best : dist);
    best = lt ? dist : best;
    bi = lt ? c : bi;
}

// Labels of the lanes in `fail` (at most 32 of the wave's lanes: points that
// failed the keep test), with the wave's lanes split over them: G = 64 /
// next_pow2(m) lanes per point, lane r of a group scanning centroids r, r+G,
// r+2G, ... with the reference arithmetic (strict '<' from FLT_MAX), then the
// group's minima combined as (distance, index) pairs, lowest index on equal
// distances -- the label the sequential scan returns (0 when no distance is
// below FLT_MAX).  Returns the label of the calling lane's own point (lanes
// outside `fail` get a value to ignore).  slot: 64 words of LDS of the wave.
// SEC: also the second minimum of the point's distances (scan_step2; the
// pairs combine with the loser's minimum joining the second ones) into sec.
template <int DMAX, int S, bool SEC = false>
__device__ __forceinline__ uint32_t sparse_scan(const float (&x)[DMAX][1], const float *__restrict__ cen_s,
                                                uint32_t d, uint32_t k, unsigned long long fail, uint32_t lane,
                                                uint32_t *slot, float *sec_out = nullptr)
{
    const uint32_t m = (uint32_t)__builtin_popcountll(fail);
    const uint32_t G = m <= 1 ? 64u : m <= 2 ? 32u : m <= 4 ? 16u : m <= 8 ? 8u : m <= 16 ? 4u : 2u;
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(fail >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)fail, 0u));
    if ((fail >> lane) & 1ull)
        slot[rank] = lane;
    mfma::wave_lds_order();
    const uint32_t j = lane / G, r = lane % G;
    const uint32_t src = j < m ? slot[j] : 0u;
    float xs[DMAX];
#pragma unroll
    for (int f = 0; f < DMAX; ++f)
        xs[f] = f < (int)d ? __shfl(x[f][0], (int)src) : 0.0f;
    float best = FLT_MAX, sec = FLT_MAX;
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t c = r; c < k; c += G) {
        const float *cc = cen_s + (size_t)c * S;
        float t = xs[0] - cc[0];
        float dist = t * t;
#pragma unroll
        for (int f = 1; f < DMAX; ++f) {
            t = xs[f] - cc[f];
            dist = __builtin_fmaf(t, t, dist);
        }
        if constexpr (SEC) {
            scan_step2(dist, c, best, bi, sec);
        } else if (dist < best) {
            best = dist;
            bi = c;
        }
    }
    for (uint32_t off = 1; off < G; off <<= 1) {
        const float ob = __shfl_xor(best, (int)off);
        const uint32_t oi = (uint32_t)__shfl_xor((int)bi, (int)off);
        const bool take = ob < best || (ob == best && oi < bi);
        if constexpr (SEC)  // symmetric: both lanes of the pair get the same
            sec = fminf(fminf(sec, __shfl_xor(sec, (int)off)), take ? best : ob);
        if (take) {
            best = ob;
            bi = oi;
        }
    }
    const uint32_t lab = bi == 0xFFFFFFFFu ? 0u : bi;
    mfma::wave_lds_order();  // slot is written again by the next call
    if constexpr (SEC)
        *sec_out = __shfl(sec, (int)(rank * G));
    return (uint32_t)__shfl((int)lab, (int)(rank * G));
}

// The sequential scan of argmin_lds (one point per lane) with the second
// minimum of scan_step2.
template <int DMAX, int S>
__device__ __forceinline__ void argmin2_lds(const float (&x)[DMAX][1], const float *__restrict__ cen,
                                            uint32_t k, uint32_t &lab, float &sec)
{
    float best = FLT_MAX;
    uint32_t bi = 0;
    sec = FLT_MAX;
    for (uint32_t c = 0; c < k; ++c) {
        const float *cc = cen + (size_t)c * S;
        float t = x[0][0] - cc[0];
        float dist = t * t;
#pragma unroll
        for (int f = 1; f < DMAX; ++f) {
            t = x[f][0] - cc[f];
            dist = __builtin_fmaf(t, t, dist);
        }
        scan_step2(dist, c, best, bi, sec);
    }
    lab = bi;
}

// ---- runs of equal labels in a wave -------------------------------------

// The direct per-lane accumulation of the plain batch tiles is taken when at
// most this many of a wave's lanes share lane 0's label and lane 32's.  On
// the 1024-job mix (one MI355X): points interleaved over the blobs 2.15 ->
// 1.87 ms per step, shuffled 2.24 -> 2.01, grouped by blob unchanged (2.23;
// direct atomics for every wave would take 3.24 there).
constexpr int kSpreadMax = 16;

constexpr uint32_t kNoLabel = 0x3FFFFFFu;  // sorts after every label (< 2^26)

// DPP controls: quad_perm [1,0,3,2] and [2,3,0,1] (lane xor 1, xor 2);
// row_shl:n / row_shr:n (lane i reads lane i+n / i-n within its row of 16);
// row_ror:8 (lane xor 8 within the row); row_bcast:15 (the last lane of each
// row to the next row) and row_bcast:31 (lane 31 to rows 2 and 3).
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppRowShl = 0x100, kDppRowShr = 0x110, kDppXor8 = 0x128,
              kDppBcast15 = 0x142, kDppBcast31 = 0x143;

template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}

// The value of lane (lane xor J): DPP within rows of 16, LDS permutes across.
template <uint32_t J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, uint32_t lane)
{
    if constexpr (J == 1)
        return dpp_move<kDppXor1, 0xF>(v);
    else if constexpr (J == 2)
        return dpp_move<kDppXor2, 0xF>(v);
    else if constexpr (J == 4) {
        const uint32_t up = dpp_move<kDppRowShl + 4, 0xF>(v), down = dpp_move<kDppRowShr + 4, 0xF>(v);
        return (lane & 4u) ? down : up;
    } else if constexpr (J == 8)
        return dpp_move<kDppXor8, 0xF>(v);
    else
        return (uint32_t)__shfl_xor((int)v, (int)J);
}

// One compare-exchange step of the bitonic network (block SIZE, stride J).
template <uint32_t SIZE, uint32_t J>
__device__ __forceinline__ uint32_t bitonic_step(uint32_t key, uint32_t lane)
{
    const uint32_t other = lane_xor<J>(key, lane);
    const bool desc = (lane & SIZE) != 0;
    const bool lower = (lane & J) == 0;
    return (lower != desc) ? min(key, other) : max(key, other);
}

template <uint32_t SIZE, uint32_t J>
__device__ __forceinline__ uint32_t bitonic_merge(uint32_t key, uint32_t lane)
{
    key = bitonic_step<SIZE, J>(key, lane);
    if constexpr (J > 1)
        key = bitonic_merge<SIZE, J / 2>(key, lane);
    return key;
}

// Bitonic sort of one 32-bit key per lane, ascending in lane order.
__device__ __forceinline__ uint32_t wave_sort(uint32_t key, uint32_t lane)
{
    key = bitonic_merge<2, 1>(key, lane);
    key = bitonic_merge<4, 2>(key, lane);
    key = bitonic_merge<8, 4>(key, lane);
    key = bitonic_merge<16, 8>(key, lane);
    key = bitonic_merge<32, 16>(key, lane);
    return bitonic_merge<64, 32>(key, lane);
}

// The wave's points sorted by label.  Lane i holds the point of lane src;
// the lanes of one label are contiguous; on the last lane of a run, last is
// set and count is the run's length.  step[s] says whether scan step s adds
// a value of the same run (steps: shifts by 1, 2, 4, 8 within rows of 16, then
// the row broadcasts of lanes 15 / 47 to rows 1 / 3 and of lane 31 to rows 2, 3).
struct WaveRuns {
    uint32_t label, src, count;
    bool last;
    bool step[6];
    uint32_t steps;  // wave-uniform: bit s set if some run needs scan step s
};

__device__ __forceinline__ WaveRuns wave_runs(uint32_t label, bool valid, uint32_t lane)
{
    const uint32_t key = wave_sort(valid ? (label << 6) | lane : 0xFFFFFFFFu, lane);
    WaveRuns r;
    r.label = key >> 6;
    r.src = key & 63u;
    // every lane takes part in every move (a lane that sits out would feed
    // the old value to its neighbours)
    const uint32_t l1 = dpp_move<kDppRowShr + 1, 0xF>(r.label);
    const uint32_t l2 = dpp_move<kDppRowShr + 2, 0xF>(r.label);
    const uint32_t l4 = dpp_move<kDppRowShr + 4, 0xF>(r.label);
    const uint32_t l8 = dpp_move<kDppRowShr + 8, 0xF>(r.label);
    const uint32_t l15 = dpp_move<kDppBcast15, 0xA>(r.label);
    const uint32_t l31 = dpp_move<kDppBcast31, 0xC>(r.label);
    const uint32_t row_lane = lane & 15u, row = lane >> 4;
    // label of lane - 1: in the row, else the previous row's last lane
    const uint32_t prev = row_lane != 0 ? l1 : (row & 1u) != 0 ? l15 : l31;
    const uint64_t heads = __ballot(lane == 0 || prev != r.label);
    const bool run_end = lane == 63 || (((heads >> lane) >> 1) & 1ull) != 0;
    r.last = run_end && r.label != kNoLabel;
    const uint64_t upto = heads & ((2ull << lane) - 1ull);  // lanes 0..lane
    r.count = lane - (63u - (uint32_t)__builtin_clzll(upto)) + 1u;
    // in-row step s (shift 2^s) matters only if a run is longer than 2^s;
    // the row broadcasts only if a run crosses a row boundary
    const uint32_t len = r.last ? r.count : 0u;
    const bool cross = r.last && (lane & 15u) + 1u < r.count;
    r.steps = (__ballot(len > 1) != 0 ? 1u : 0u) | (__ballot(len > 2) != 0 ? 2u : 0u) |
              (__ballot(len > 4) != 0 ? 4u : 0u) | (__ballot(len > 8) != 0 ? 8u : 0u) |
              (__ballot(cross) != 0 ? 48u : 0u);
    r.step[0] = (row_lane >= 1) & (l1 == r.label);
    r.step[1] = (row_lane >= 2) & (l2 == r.label);
    r.step[2] = (row_lane >= 4) & (l4 == r.label);
    r.step[3] = (row_lane >= 8) & (l8 == r.label);
    r.step[4] = ((row & 1u) != 0) & (l15 == r.label);
    r.step[5] = (row >= 2) & (l31 == r.label);
    return r;
}

// Inclusive scan of v along the sorted lanes within runs: the last lane of a
// run ends with the run's sum (exact: at most 64 grid integers, |sum| < 2^31).
// Steps that no run needs are skipped (wave-uniform branches): a step adds
// only within a run, and a run of length L is complete after the in-row
// shifts below L plus, if it crosses a row boundary, the broadcasts.
__device__ __forceinline__ int32_t run_sum(int32_t v, const WaveRuns &r)
{
    int32_t t;
    if (r.steps & 1u) {
        t = (int32_t)dpp_move<kDppRowShr + 1, 0xF>((uint32_t)v);
        v = r.step[0] ? v + t : v;
    }
    if (r.steps & 2u) {
        t = (int32_t)dpp_move<kDppRowShr + 2, 0xF>((uint32_t)v);
        v = r.step[1] ? v + t : v;
    }
    if (r.steps & 4u) {
        t = (int32_t)dpp_move<kDppRowShr + 4, 0xF>((uint32_t)v);
        v = r.step[2] ? v + t : v;
    }
    if (r.steps & 8u) {
        t = (int32_t)dpp_move<kDppRowShr + 8, 0xF>((uint32_t)v);
        v = r.step[3] ? v + t : v;
    }
    if (r.steps & 16u) {
        t = (int32_t)dpp_move<kDppBcast15, 0xA>((uint32_t)v);
        v = r.step[4] ? v + t : v;
        t = (int32_t)dpp_move<kDppBcast31, 0xC>((uint32_t)v);
        v = r.step[5] ? v + t : v;
    }
    return v;
}

// Partial row of a tile from the exact accumulator: sums [d][k] (grid
// integers times the unit, exact in FP64), then the counts [k].
__device__ __forceinline__ void write_partial_row(const BAcc &a, uint32_t d, uint32_t k, double *prow)
{
    for (uint32_t i = threadIdx.x; i < k * (d + 1); i += blockDim.x)
        prow[i] = i < d * k ? (double)a.A[i] * grid_unit(a.scl[i / k]) : (double)a.cnt[i - d * k];
}

template <int DMAX, bool FULL>
__device__ __forceinline__ void batch_tile(const BJob &J, uint32_t tindex, float *lds, uint32_t waves)
{
    (void)waves;
    const uint32_t tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const uint32_t d = J.d, k = J.k;
    constexpr int S = (int)cen_stride(DMAX);
    const BAcc acc = acc_carve(reinterpret_cast<char *>(lds), J, DMAX);
    float *cen_s = reinterpret_cast<float *>(reinterpret_cast<char *>(lds) + acc_bytes(k, k * d, DMAX));
    uint32_t *slot = reinterpret_cast<uint32_t *>(cen_s + (size_t)k * S) + wave * 64;

    stage_centroids<DMAX, S>(J.cen[J.iter & 1], d, k, cen_s);
    __syncthreads();

    const uint64_t first = (uint64_t)tindex * J.tile_pts;
    const uint64_t end = std::min<uint64_t>(first + J.tile_pts, J.n);
    const uint64_t n = J.n;
    // The job's arrays as global-memory pointers, read from the record once
    // (through the record they would be flat loads, each waiting for every
    // earlier load).
    const gfloat *pts = (const gfloat *)J.pts;
    const guint *prev_labels = (const guint *)J.labels;
    const gfloat *keep_thr = (const gfloat *)J.keep;
    // The next pass's points and previous labels are loaded while this pass
    // is labeled.  Every load is unconditional and in bounds (past the tile:
    // its first point; past d: feature d-1, zeroed at use), so no branch
    // separates the loads from their use.
    const bool keep_ok = J.keep_ok != 0;
    const bool special = J.gmax[d] != 0;  // the job holds NaN / inf
    // d <= 8: the second-distance bound (second_bound), kept for every point
    constexpr bool kBound = DMAX <= 8;
    const bool bound_ok = kBound && J.bound_ok != 0;
    const gfloat *lb_in = (const gfloat *)J.lb;
    float *lb_out = J.lb;
    const float drift = J.drift;
    const float shrink = drift == 0.0f ? 1.0f : kBoundShrink;  // B - 0 is exact
    float xn[DMAX];
    uint32_t ln = 0;
    float bn = 0.0f;
    // byte offsets into the points fit 32 bits (n*d <= 2^30, checked at
    // admission): one base, one 32-bit offset advanced by a row per feature
    const uint32_t row_bytes = (uint32_t)n * 4u;
    // The feature count as the loop body sees it.  At DMAX = 32 it is made
    // opaque at every pass (an empty asm), so the per-feature tests (f < d)
    // are evaluated where they are used instead of being hoisted out of the
    // loop as 32 masks that crowd the scalar registers and spill into vector
    // ones (smaller buckets keep their hoisted masks: they fit).  FULL
    // (d == DMAX): every feature is real and the tests vanish.
    uint32_t dq = FULL ? (uint32_t)DMAX : d;
    auto fetch = [&](uint64_t b0) {
        const uint64_t p = b0 + tid;
        const uint64_t pc = p < end ? p : first;
        uint32_t off = (uint32_t)pc * 4u;
#pragma unroll
        for (int f = 0; f < DMAX; ++f) {
            xn[f] = *(const gfloat *)((const __attribute__((address_space(1))) char *)pts + off);
            // next row; past d: feature d-1 again, zeroed at use (the step
            // is wave-uniform: a scalar select)
            off += (f + 1 < (int)dq) ? row_bytes : 0u;
        }
        // previous labels only matter to jobs with the keep test; the
        // others read one cached word instead of streaming the labels
        ln = prev_labels[keep_ok || bound_ok ? pc : first];
        if constexpr (kBound)
            bn = lb_in[bound_ok ? pc : first];
    };
    fetch(first);
    for (uint64_t base = first; base < end; base += blockDim.x) {
        const uint64_t p = base + tid;
        const bool valid = p < end;
        if constexpr (DMAX >= 32 && !FULL)
            asm volatile("" : "+s"(dq));
        float x[DMAX][1];
#pragma unroll
        for (int f = 0; f < DMAX; ++f)
            x[f][0] = (f < (int)dq) ? xn[f] : 0.0f;
        const uint32_t prev = ln;
        const uint32_t a = prev < k ? prev : 0u;
        const float bprev = bn;
        // the keep threshold is loaded before the next pass's points, so
        // waiting for it does not wait for them
        const float thr = keep_ok ? keep_thr[a] : -1.0f;
        fetch(base + blockDim.x);
        // a point inside its previous centroid's keep radius keeps the label
        // (the scan would return it, keep_threshold), and so does one closer
        // to it than its bound allows any other centroid (second_bound); the
        // others scan
        uint32_t lab[1] = {0u};
        bool keep = false;
        float bkeep = 0.0f;  // a kept point's bound (0 when none is valid)
        if (keep_ok || bound_ok) {
            const float *ca = cen_s + (size_t)a * S;
            float t = x[0][0] - ca[0];
            float dist = t * t;
#pragma unroll
            for (int f = 1; f < DMAX; ++f) {
                t = x[f][0] - ca[f];
                dist = __builtin_fmaf(t, t, dist);
            }
            float lsq = 0.0f;
            if (bound_ok) {
                bkeep = fmaxf((bprev - drift) * shrink, 0.0f);
                lsq = bkeep * bkeep * kBoundMargin;
                lsq = lsq >= kBoundMin ? lsq : 0.0f;
            }
            keep = prev < k && (dist <= thr || dist < lsq);
            lab[0] = a;
        }
        const unsigned long long fail = __ballot(valid && !keep);
        float sec = FLT_MAX;
        if (fail != 0) {
            // few failures: the whole wave rescans those points; many: each
            // failing lane scans its own
            if (__builtin_popcountll(fail) <= 32) {
                const uint32_t l = sparse_scan<DMAX, S, kBound>(x, cen_s, dq, k, fail, lane, slot, &sec);
                if (valid && !keep)
                    lab[0] = l;
            } else if (valid && !keep) {
                if constexpr (kBound)
                    argmin2_lds<DMAX, S>(x, cen_s, k, lab[0], sec);
                else
                    argmin_lds<DMAX, 1, S>(x, cen_s, k, lab);
            }
            if (valid && !keep)
                J.labels[p] = lab[0];
        }
        if constexpr (kBound) {
            // every point's bound for the next iteration (a kept one's is
            // unchanged when nothing moved)
            if (valid && !(keep && bound_ok && drift == 0.0f))
                lb_out[p] = keep ? bkeep : second_bound(sec);
        }
        // grid integers into the work-group's accumulator (a job without
        // NaN / inf is on its grid: no range checks)
        bool off = false;
        auto qv = [&](int f) {
            if (special) {
                bool ok;
                const int32_t q = grid_q(x[f][0], acc.scl[f], ok);
                off = off || (valid && !ok);
                return q;
            }
            return grid_q_on(x[f][0], acc.scl[f]);
        };
        // Labels spread over the wave (at most kSpreadMax lanes share lane
        // 0's or lane 32's label): every lane adds its own values with LDS
        // integer atomics.  Otherwise (long runs of one label: data grouped
        // by cluster) the lanes are sorted by label and each run's last lane
        // adds the run's sums.  Either way the integer sums are exact, so
        // the two forms give the same bits.
        const uint32_t l0 = __builtin_amdgcn_readfirstlane(lab[0]);
        const uint32_t l32 = __builtin_amdgcn_readlane(lab[0], 32);
        const bool spread = __builtin_popcountll(__ballot(valid && lab[0] == l0)) <= kSpreadMax &&
                            __builtin_popcountll(__ballot(valid && lab[0] == l32)) <= kSpreadMax;
        if (spread) {
#pragma unroll
            for (int f = 0; f < DMAX; ++f)
                if (f < (int)dq) {
                    const int32_t q = qv(f);
                    if (valid)
                        lds_add64(acc.A + (uint32_t)f * k + lab[0], (long long)q);
                }
            if (valid)
                lds_add(acc.cnt + lab[0], 1u);
        } else {
            const WaveRuns r = wave_runs(lab[0], valid, lane);
#pragma unroll
            for (int f = 0; f < DMAX; ++f)
                if (f < (int)dq) {
                    const int32_t v = run_sum((int32_t)__shfl((int)qv(f), (int)r.src), r);
                    if (r.last)
                        lds_add64(acc.A + (uint32_t)f * k + r.label, (long long)v);
                }
            if (r.last)
                lds_add(acc.cnt + r.label, r.count);
        }
        if (special && __builtin_expect(__ballot(off) != 0, 0)) {
            // off-grid values (added as 0): flag their cells for the finish
            if (off) {
#pragma unroll
                for (int f = 0; f < DMAX; ++f)
                    if (f < (int)dq && !(__builtin_fabsf(x[f][0] * acc.scl[f]) < kGridLimit))
                        atomicOr(&J.flags[(size_t)f * k + lab[0]], special_flag(x[f][0]));
            }
        }
        mfma::wave_lds_order();
    }
    __syncthreads();
    write_partial_row(acc, d, k, J.part + (size_t)tindex * ((size_t)k * (d + 1)));
}

// Lanes per element in the merge of a job's tile rows: a power of two <= 64
// that depends only on the row length and the tile count (never on the
// launch), so a job's merged sums are the same in every batch.
__host__ __device__ inline uint32_t merge_split(uint32_t row, uint32_t nt)
{
    uint32_t s = 1;
    while (s < 64 && 2 * s <= nt && (2 * s) * row <= 64 * kBatchMaxWaves)
        s *= 2;
    return s;
}

constexpr uint32_t kMergeThreads = 256;

// A merge work-group's share: elements [e0, e0 + kMergeThreads / S) of one
// job's row.
struct BMergeUnit {
    uint32_t job, e0;
};

// Merge a job's tile rows for one element slice.  Element e of the row (sums
// [d][k], then masses [k]) is summed by S lanes of one wave: lane s adds rows
// s, s+S, s+2S, ... in order, then the S lane sums combine by an xor
// butterfly (every lane ends with the same bits); the merged value overwrites
// row 0 (read only by the lane that writes it), where the finish reads it.
__device__ void merge_rows(const BJob &J, uint32_t e0)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t nt = J.ntiles, rl = J.k * (J.d + 1);
    const uint32_t S = merge_split(rl, nt);
    const uint32_t per = kMergeThreads / S, s = tid % S;
    const uint32_t e = e0 + tid / S;
    const bool mine = e < rl && e < e0 + per;
    double acc = 0.0;
    if (mine) {
        const gdouble *col = (const gdouble *)(J.part + e);  // global, not flat, loads
        uint32_t t = s;
        for (; t + 3 * S < nt; t += 4 * S) {
            const double a0 = col[(size_t)t * rl], a1 = col[(size_t)(t + S) * rl];
            const double a2 = col[(size_t)(t + 2 * S) * rl], a3 = col[(size_t)(t + 3 * S) * rl];
            acc += a0;
            acc += a1;
            acc += a2;
            acc += a3;
        }
        for (; t < nt; t += S)
            acc += col[(size_t)t * rl];
    }
    for (uint32_t o = 1; o < S; o *= 2)
        acc += __shfl_xor(acc, (int)o);
    if (mine && s == 0)
        J.part[e] = acc;
}

// Divide the merged row 0, shift, advance the job.
// keep_threshold of every centroid of [d][k] (global) into keep[k]: FP64
// squared distances to the other centroids, NaN distances skipped.  The
// work-group stages the centroids in LDS as rows [k][DB] zero-padded to the
// bucket DB >= d (padding adds 0 exactly) when they fit; each wave takes 64
// centroids a (one per lane, its row in registers) against a slice of the
// centroids c, unrolled for independent LDS reads, and the slices' minima
// meet in red.  The minimum does not depend on the split.
constexpr uint32_t kKeepStage = 12288;  // floats of LDS for the staged centroids

// Width of a staged centroid row: d rounded up to 4 / 8 / 16 / 32; 0 beyond.
__host__ __device__ inline uint32_t keep_row_width(uint32_t d)
{
    return d <= 4 ? 4u : d <= 8 ? 8u : d <= 16 ? 16u : d <= 32 ? 32u : 0u;
}
// FP32-scan jobs with fewer clusters and k*d below kKeepMinKd scan every
// centroid: their scan (about 2*k*d VALU ops per point) is too short to
// repay the keep test (2*d ops, a threshold and the previous label read).
constexpr uint32_t kKeepMinK = 16;
constexpr uint32_t kKeepMinKd = 96;

template <int DB>
__device__ __forceinline__ double keep_min_rows(const float *__restrict__ cs, uint32_t a, uint32_t c0, uint32_t c1)
{
    double ca[DB];
#pragma unroll
    for (int f = 0; f < DB; ++f)
        ca[f] = (double)cs[(size_t)a * DB + f];
    // (c unrolled by U: independent LDS reads in flight, within the registers)
    constexpr int U = DB <= 8 ? 4 : DB == 16 ? 2 : 1;
    double s2 = __builtin_inf();
    uint32_t c = c0;
    for (; c + U <= c1; c += U) {
        double acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc[u] = 0.0;
#pragma unroll
            for (int f = 0; f < DB; ++f) {
                const double t = ca[f] - (double)cs[(size_t)(c + u) * DB + f];
                acc[u] = __builtin_fma(t, t, acc[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            s2 = (c + u != a && acc[u] < s2) ? acc[u] : s2;  // a NaN distance never lowers it
    }
    for (; c < c1; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int f = 0; f < DB; ++f) {
            const double t = ca[f] - (double)cs[(size_t)c * DB + f];
            acc = __builtin_fma(t, t, acc);
        }
        s2 = (c != a && acc < s2) ? acc : s2;
    }
    return s2;
}

// rows_out (optional): the staged rows also written to global memory.
// thresholds = false: only the rows.
__device__ __forceinline__ void build_keep_thresholds(const float *__restrict__ cen, uint32_t d, uint32_t k,
                                      float *__restrict__ keep, float *cs, double *red,
                                      float *__restrict__ rows_out = nullptr, bool thresholds = true)
{
    const uint32_t tid = threadIdx.x, lane = tid % 64, wave = tid / 64, W = blockDim.x / 64;
    const uint32_t db = keep_row_width(d);
    const bool staged = db != 0 && (uint64_t)k * db <= kKeepStage;
    if (staged)
        for (uint32_t i = tid; i < k * db; i += blockDim.x) {
            const uint32_t c = i / db, f = i % db;
            cs[i] = f < d ? cen[(size_t)f * k + c] : 0.0f;
            if (rows_out)
                rows_out[i] = cs[i];
        }
    __syncthreads();
    if (!thresholds)
        return;
    const uint32_t nab = (k + 63) / 64;
    const uint32_t slices = nab >= W ? 1u : W / nab;
    for (uint32_t ab0 = 0; ab0 < nab; ab0 += W / slices) {
        const uint32_t ab = ab0 + wave / slices, sl = wave % slices;
        const uint32_t a = ab * 64 + lane;
        const bool mine = ab < nab && a < k;
        double s2 = __builtin_inf();
        if (mine) {
            const uint32_t per = (k + slices - 1) / slices;
            const uint32_t c0 = sl * per, c1 = min(k, c0 + per);
            if (staged) {
                switch (db) {
                case 4: s2 = keep_min_rows<4>(cs, a, c0, c1); break;
                case 8: s2 = keep_min_rows<8>(cs, a, c0, c1); break;
                case 16: s2 = keep_min_rows<16>(cs, a, c0, c1); break;
                default: s2 = keep_min_rows<32>(cs, a, c0, c1); break;
                }
            } else {
                for (uint32_t c = c0; c < c1; ++c) {
                    double acc = 0.0;
                    for (uint32_t f = 0; f < d; ++f) {
                        const double t = (double)cen[(size_t)f * k + a] - (double)cen[(size_t)f * k + c];
                        acc = __builtin_fma(t, t, acc);
                    }
                    s2 = (c != a && acc < s2) ? acc : s2;
                }
            }
        }
        if (slices > 1) {
            red[tid] = s2;
            __syncthreads();
            if (sl == 0 && mine)
                for (uint32_t q = 1; q < slices; ++q)
                    s2 = fmin(s2, red[tid + q * 64]);
            __syncthreads();
        }
        if (sl == 0 && mine)
            keep[a] = keep_threshold(s2);
    }
}

__device__ void batch_finalize(BJob &J, double *red)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t d = J.d, k = J.k;
    const double *tot = J.part;
    const float *old = J.cen[J.iter & 1];
    float *nxt = J.cen[(J.iter & 1) ^ 1];
    double local = 0.0, moved = 0.0;
    const bool special = J.gmax[d] != 0;  // the job holds NaN / inf: flags may be set
    for (uint32_t c = tid; c < k; c += blockDim.x) {
        const double m = tot[(size_t)k * d + c];
        double sq = 0.0;
        bool fin_new = true, fin_old = true;
        for (uint32_t f = 0; f < d; ++f) {
            // exact sums, or NaN / inf where off-grid values were flagged
            // (the flags are cleared for the next iteration)
            const size_t i = (size_t)f * k + c;
            const uint32_t fl = special ? J.flags[i] : 0u;
            if (fl)
                J.flags[i] = 0u;
            const float v = (float)((fl ? special_value(fl) : tot[i]) / m);
            const float o = old[(size_t)f * k + c];
            nxt[(size_t)f * k + c] = v;
            const double diff = (double)v - (double)o;
            sq += diff * diff;
            fin_new = fin_new && __builtin_isfinite(v);
            fin_old = fin_old && __builtin_isfinite(o);
        }
        J.masses[c] = (uint32_t)m;
        if (m != 0.0)
            local = fmax(local, sqrt(sq));
        // the drift bound covers the centroids that can be chosen next
        // (finite ones); one that was not finite before is unbounded
        if (fin_new)
            moved = fmax(moved, fin_old ? sqrt(sq) : __builtin_inf());
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
        if (J.iter >= J.max_iter || (J.tol > 0.0 && red[0] < J.tol))
            J.state = CLK_JOB_DONE;
    }
    if (J.lb == nullptr)
        return;
    // drift: the largest move, FP64 from the FP32 coordinates (a few 2^-53
    // relative), widened by 2^-40 and rounded up to FP32
    __syncthreads();
    red[tid] = moved;
    __syncthreads();
    if (tid < blockDim.x - p2)
        red[tid] = fmax(red[tid], red[tid + p2]);
    __syncthreads();
    for (uint32_t s = p2 / 2; s > 0; s /= 2) {
        if (tid < s)
            red[tid] = fmax(red[tid], red[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        J.drift = __double2float_ru(red[0] * (1.0 + 0x1p-40));
        J.bound_ok = 1;  // the tiles wrote every point's bound in this iteration
    }
}

// The persistent tile loop: work-group-uniform item numbers from the global
// counter, the next one requested before the current tile is processed.
template <typename Body>
__device__ __forceinline__ void for_each_tile(const BJob *__restrict__ jobs, const BTile *__restrict__ tiles,
                                              BCtl *__restrict__ ctl, Body body)
{
    __shared__ uint32_t s_item;
    const uint32_t tid = threadIdx.x;
    const uint32_t ntiles = ctl->ntiles;
    uint32_t next = 0;
    if (tid == 0)
        next = atomicAdd(&ctl->next, 1u);
    for (;;) {
        if (tid == 0)
            s_item = next;
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();  // s_item is written again at the next loop head
        if (item >= ntiles)
            break;
        if (tid == 0)
            next = atomicAdd(&ctl->next, 1u);
        const BTile t = tiles[item];
        const BJob &J = jobs[t.job];
        if (J.state == CLK_JOB_ACTIVE)
            body(J, t.index);
    }
}

// Tiles of the jobs of one size class (FP32 labeling, d <= DMAX).
// d <= 32: at most 168 VGPRs, so three work-groups fit per CU (the
// prefetched and the current point's 32 features are both live).
template <int DMAX>
__global__ void __launch_bounds__(64 * kBatchMaxWaves, DMAX == 32 ? 3 : 1)
batch_tiles_kernel(const BJob *__restrict__ jobs, const BTile *__restrict__ tiles, BCtl *__restrict__ ctl)
{
    extern __shared__ __align__(16) float lds[];
    const uint32_t waves = blockDim.x / 64;
    for_each_tile(jobs, tiles, ctl, [&](const BJob &J, uint32_t index) {
        if (J.d == DMAX)
            batch_tile<DMAX, true>(J, index, lds, waves);
        else
            batch_tile<DMAX, false>(J, index, lds, waves);
        __syncthreads();  // the tables are cleared for the next tile
    });
}

// Read in place of a missing point or feature (the loads need no masks).
__device__ const float kZeroWord = 0.0f;

template <bool FP16, int NKB, bool FULL>
__device__ __forceinline__ void batch_tile_mfma(const BJob &J, uint32_t tindex, char *lds, const BLayout &Lo)
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
    const BAcc acc = acc_carve(lds + Lo.o_acc, J, P);
    uint32_t *pl = reinterpret_cast<uint32_t *>(lds + Lo.o_pairs) + (size_t)wave * kBatchPairs;
    unsigned long long *wmin = reinterpret_cast<unsigned long long *>(lds + Lo.o_min) + (size_t)wave * 64;
    uint32_t *pc = reinterpret_cast<uint32_t *>(lds + Lo.o_count) + wave;
    const float *cen = J.cen[J.iter & 1];
    // keep test (jobs with centroid rows): a point inside its previous
    // centroid's keep radius keeps its label (keep_threshold); a wave whose
    // points all pass skips the scores, one with one or two failures rescans
    // those exactly from the rows, others take the MFMA path
    const bool special = J.gmax[d] != 0;  // the job holds NaN / inf
    const uint32_t w4 = J.cen_t_w / 4;
    // global-memory views of the job's arrays (global, not flat, loads)
    const gf32v4 *rows = (const gf32v4 *)J.cen_t;
    const gfloat *gpts = (const gfloat *)J.pts;
    const guint *glabels = (const guint *)J.labels;
    const gfloat *gkeep = (const gfloat *)J.keep;

    for (uint32_t e = tid; e < nt16 * NKB * 64; e += blockDim.x)
        ltab[e] = J.table[e];
    for (uint32_t c = tid; c < nt16 * kTile; c += blockDim.x)
        lnorm[c] = J.norm[c];
    __syncthreads();
    const bool table_ok = J.stats[2] == 0.0f;
    const uint32_t kmask = (1u << J.prm.key_bits) - 1u;
    const uint64_t first = (uint64_t)tindex * J.tile_pts;
    const uint64_t end = std::min<uint64_t>(first + J.tile_pts, n);
    const uint64_t nb = (end - first + 63) / 64;
    // The points are loaded one point per lane (xc: keep test, exact rescans,
    // accumulation).  The MFMA lane layout (raw) is needed only by batches
    // that take the MFMA path: it is prefetched for the next batch while
    // this one is labeled when this one took that path (or the job has no
    // keep test yet), else loaded on demand -- once the labels settle, most
    // batches pass the keep test and read each point once.
    auto batch_ids = [&](uint64_t b0, uint64_t (&pid)[kGroups]) {
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const uint64_t p = b0 + g * 16 + q;
            pid[g] = p < end ? p : ~0ull;
        }
    };
    float raw[kGroups][NKB * 8];
    auto fetch = [&](uint64_t b0) {
        uint64_t pid[kGroups];
        batch_ids(b0, pid);
        fetch_points<NKB>(raw, J.pts, n, FULL ? P : d, pid, h, &kZeroWord);
    };
    const bool keep_ok = J.keep_ok != 0;
    bool pre = !keep_ok && wave < nb;  // raw holds this batch's points
    if (pre)
        fetch(first + (uint64_t)wave * 64);
    // Rounds: in round rd wave w labels batch rd*W + w; the run sums go to
    // the work-group's accumulator (exact integer atomics, any order).
    const uint32_t rounds = (uint32_t)((nb + waves - 1) / waves);
    // d as the rounds see it, opaque at every round (an empty asm): the
    // per-feature tests are evaluated where they are used, not hoisted out
    // of the loop as masks that crowd the scalar registers.  FULL (d equal
    // to the K-blocks' 32 * NKB features): the tests vanish.
    uint32_t dq = FULL ? (uint32_t)(32 * NKB) : d;
    for (uint32_t rd = 0; rd < rounds; ++rd) {
        if constexpr (!FULL)
            asm volatile("" : "+s"(dq));
        const uint64_t bt = (uint64_t)rd * waves + wave;
        WaveRuns r;
        r.last = false;
        r.label = 0;
        r.count = 0;
        float xc[P];
        if (bt < nb) {
        const uint64_t base = first + bt * 64;
        uint64_t pid[kGroups];
        batch_ids(base, pid);
        const uint64_t pme = base + lane;  // lane l holds point base + l in xc
        const bool vme = pme < end;
        uint32_t prev = 0;
        {
            // unconditional in-bounds loads (past d: feature d-1, zeroed)
            const uint64_t pc = vme ? pme : first;
            const uint32_t off = (uint32_t)pc * 4u;  // 32-bit byte offsets (n*d <= 2^30)
            const uint32_t row_bytes = (uint32_t)n * 4u;
#pragma unroll
            for (int f = 0; f < (int)P; ++f) {
                const uint32_t row = (uint32_t)min(f, (int)dq - 1) * row_bytes;
                const float v = *(const gfloat *)((const __attribute__((address_space(1))) char *)gpts + (off + row));
                xc[f] = f < (int)dq ? v : 0.0f;
            }
            prev = glabels[keep_ok ? pc : first];  // one cached word without the keep test
        }
        // The keep test runs before the next batch's loads are issued: loads
        // complete in order, so a load issued after them (the centroid row,
        // the threshold) would wait for them.
        bool keep = false;
        if (keep_ok) {
            const uint32_t a = prev < k ? prev : 0u;
            float dist = 0.0f;
#pragma unroll
            for (int q = 0; q < (int)P / 4; ++q)
                if ((uint32_t)q < w4) {
                    const f32v4 c4 = rows[(size_t)a * w4 + q];
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (4 * q + j < (int)dq) {
                            const float t = xc[4 * q + j] - c4[j];
                            dist = __builtin_fmaf(t, t, dist);
                        }
                }
            keep = prev < k && dist <= gkeep[a];
        }
        const unsigned long long fail = keep_ok ? __ballot(vme && !keep) : ~0ull;
        const uint32_t nfail = (uint32_t)__builtin_popcountll(fail);
        const bool scores = !(keep_ok && nfail <= 2);  // wave-uniform
        PointOps<NKB> x;
        float X2[kGroups];
        if (scores) {
            if (!pre)
                fetch(base);
            build_operands<FP16, NKB>(raw, x, X2);
        }
        pre = scores;
        if (scores)
            fetch(first + (bt + waves) * 64);  // past the tile: every point missing, zeros read
        uint32_t mine;
        uint64_t mp;
        if (!scores) {
            // kept labels stand; one or two failing points are rescanned
            // exactly: G lanes per point over centroids r, r+G, ..., the
            // (distance, index) minima combined with the lowest index on
            // ties (the reference's strict '<' scan)
            mine = prev < k ? prev : 0u;
            mp = vme ? pme : ~0ull;
            if (nfail != 0) {
                const uint32_t G = nfail == 1 ? 64u : 32u;
                const uint32_t src0 = (uint32_t)__builtin_ctzll(fail);
                const uint32_t src1 = nfail == 2 ? (uint32_t)__builtin_ctzll(fail & (fail - 1)) : src0;
                const uint32_t grp = lane / G, r = lane % G;
                const uint32_t src = grp == 0 ? src0 : src1;
                float xs[P];
#pragma unroll
                for (int f = 0; f < (int)P; ++f)
                    xs[f] = f < (int)dq ? __shfl(xc[f], (int)src) : 0.0f;
                float best = FLT_MAX;
                uint32_t bi = 0xFFFFFFFFu;
                for (uint32_t c = r; c < k; c += G) {
                    float dist = 0.0f;
#pragma unroll
                    for (int q = 0; q < (int)P / 4; ++q)
                        if ((uint32_t)q < w4) {
                            const f32v4 c4 = rows[(size_t)c * w4 + q];
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (4 * q + j < (int)dq) {
                                    const float t = xs[4 * q + j] - c4[j];
                                    dist = __builtin_fmaf(t, t, dist);
                                }
                        }
                    if (dist < best) {
                        best = dist;
                        bi = c;
                    }
                }
                for (uint32_t off = 1; off < G; off <<= 1) {
                    const float ob = __shfl_xor(best, (int)off);
                    const uint32_t oi = (uint32_t)__shfl_xor((int)bi, (int)off);
                    if (ob < best || (ob == best && oi < bi)) {
                        best = ob;
                        bi = oi;
                    }
                }
                const uint32_t lab = bi == 0xFFFFFFFFu ? 0u : bi;
                // the failing lane takes its group's result
                const uint32_t got0 = (uint32_t)__shfl((int)lab, 0);
                const uint32_t got1 = (uint32_t)__shfl((int)lab, (int)G % 64);
                if (vme && !keep) {
                    mine = lane == src0 ? got0 : got1;
                    if (mine != prev)
                        J.labels[pme] = mine;
                }
            }
        } else {
        const WaveBounds w = wave_bounds<NKB, FP16>(X2, J.stats, J.prm);
        uint32_t best[kGroups], second[kGroups];
        rank_all<FP16, NKB>(ltab, lnorm, nt16, lane, h, w.K, kmask, x, best, second);
        MarginDecision v[kGroups];
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
        // labels: lane (q, h) stores group h's point (= base + lane)
        mine = lab[0];
        mp = pid[0];
#pragma unroll
        for (int g = 1; g < kGroups; ++g)
            if ((uint32_t)g == h) {
                mine = lab[g];
                mp = pid[g];
            }
        if (mp < n)
            J.labels[mp] = mine;
        }
        // lane l holds point base + l (xc): its coordinates become grid
        // integers (kept in xc's registers), added lane by lane when the
        // labels are spread over the wave, else summed along the runs of the
        // points sorted by label (as in the plain tiles)
        const bool valid = mp < n;
        const uint32_t l0 = __builtin_amdgcn_readfirstlane(mine);
        const uint32_t l32 = __builtin_amdgcn_readlane(mine, 32);
        const bool spread = __builtin_popcountll(__ballot(valid && mine == l0)) <= kSpreadMax &&
                            __builtin_popcountll(__ballot(valid && mine == l32)) <= kSpreadMax;
        bool off = false;
        if (special) {
#pragma unroll
            for (int f = 0; f < (int)P; ++f)
                if (f < (int)dq) {
                    bool ok;
                    xc[f] = __int_as_float(grid_q(xc[f], acc.scl[f], ok));
                    off = off || (mp < n && !ok);
                }
        } else {
#pragma unroll
            for (int f = 0; f < (int)P; ++f)
                if (f < (int)dq)
                    xc[f] = __int_as_float(grid_q_on(xc[f], acc.scl[f]));
        }
        if (special && __builtin_expect(__ballot(off) != 0, 0)) {
            // off-grid values (added as 0): flag their cells for the finish
            if (off)
                for (uint32_t f = 0; f < d; ++f) {
                    const float v = gpts[(uint64_t)f * n + mp];
                    if (!(__builtin_fabsf(v * acc.scl[f]) < kGridLimit))
                        atomicOr(&J.flags[(size_t)f * k + mine], special_flag(v));
                }
        }
        if (spread) {
            if (valid) {
#pragma unroll
                for (int f = 0; f < (int)P; ++f)
                    if (f < (int)dq)
                        lds_add64(acc.A + (uint32_t)f * k + mine, (long long)__float_as_int(xc[f]));
                lds_add(acc.cnt + mine, 1u);
            }
        } else {
            r = wave_runs(mine, valid, lane);
#pragma unroll
            for (int f = 0; f < (int)P; ++f)
                if (f < (int)dq)
                    xc[f] = __int_as_float(run_sum((int32_t)__shfl(__float_as_int(xc[f]), (int)r.src), r));
        }
        }
        if (r.last) {
#pragma unroll
            for (int f = 0; f < (int)P; ++f)
                if (f < (int)dq)
                    lds_add64(acc.A + (uint32_t)f * k + r.label, (long long)__float_as_int(xc[f]));
            lds_add(acc.cnt + r.label, r.count);
        }
        wave_lds_order();
    }
    __syncthreads();
    write_partial_row(acc, d, k, J.part + (size_t)tindex * ((size_t)k * (d + 1)));
}

// Tiles of the jobs of one size class (MFMA labeling, NKB K-blocks).
template <bool FP16, int NKB>
__global__ void __launch_bounds__(NKB == 1 ? 64 * kBatchMfmaWaves : 64 * kBatchMaxWaves)
batch_tiles_mfma_kernel(const BJob *__restrict__ jobs, const BTile *__restrict__ tiles, BCtl *__restrict__ ctl,
                        BLayout Lo)
{
    extern __shared__ __align__(16) char lds_b[];
    for_each_tile(jobs, tiles, ctl, [&](const BJob &J, uint32_t index) {
        if (J.d == 32 * NKB)
            batch_tile_mfma<FP16, NKB, true>(J, index, lds_b, Lo);
        else
            batch_tile_mfma<FP16, NKB, false>(J, index, lds_b, Lo);
        __syncthreads();
    });
}

// Size classes: FP32 scan, d bucket 3 / 4 / 8 / 16 / 32 / 64 (classes 0-11);
// MFMA, d <= 16 / d <= 32 / d <= 64 (12-17); each with k <= 64 (even) or
// k > 64 (odd).
constexpr uint32_t kClasses = 18;
constexpr uint32_t kMfmaClass0 = 12, kMfmaNkb2Class = 16;
// Register buckets of the FP32 classes (a padded feature costs a sub and an
// fma per centroid, so d = 3 gets its own bucket).
constexpr uint32_t kFp32Dmax[] = {3, 4, 8, 16, 32, 64};

// New tilings of jobs (slot, tile_pts, ntiles triples) into their records.
__global__ void __launch_bounds__(256)
batch_retile_kernel(BJob *__restrict__ jobs, const uint32_t *__restrict__ upd, uint32_t count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    BJob &J = jobs[upd[3 * i]];
    J.tile_pts = upd[3 * i + 1];
    J.ntiles = upd[3 * i + 2];
}

// The merge of every active job's tile rows: one work-group per element
// slice of a job (large jobs get many).  It runs in the autotuner's dry
// passes too (its cost depends on the tiling); the merged row 0 it leaves
// is rewritten by the next iteration's tiles.
__global__ void __launch_bounds__(kMergeThreads)
batch_merge_kernel(const BJob *__restrict__ jobs, const BMergeUnit *__restrict__ units, const BCtl *__restrict__ ctl)
{
    (void)ctl;
    const BMergeUnit u = units[blockIdx.x];
    const BJob &J = jobs[u.job];
    if (J.state != CLK_JOB_ACTIVE)
        return;
    merge_rows(J, u.e0);
}

// One work-group per active job: divide, advance, and for a job with the MFMA
// labeling the next iteration's 16-bit table.  Work-group 0 resets the tile
// counters of every class (the tile kernels have finished).
// The status fields that lead a job record (iter, state, shift), mirrored
// by the finish kernel into pinned host memory after every iteration.
struct BStatus {
    uint32_t iter, state;
    double shift;
};

template <bool FP16>
__global__ void __launch_bounds__(256)
batch_finish_kernel(BJob *__restrict__ jobs, const uint32_t *__restrict__ active, BCtl *__restrict__ ctl,
                    BStatus *__restrict__ status)
{
    __shared__ double red[256];
    __shared__ float cs[kKeepStage];
    if (blockIdx.x == 0 && threadIdx.x < kClasses)
        ctl[threadIdx.x].next = 0;
    if (ctl[0].dry != 0)
        return;
    BJob &J = jobs[active[blockIdx.x]];
    if (J.state != CLK_JOB_ACTIVE)
        return;
    const uint32_t nxt = (J.iter & 1) ^ 1;
    const bool table = J.kind == 1;
    batch_finalize(J, red);
    __syncthreads();
    if (threadIdx.x == 0)  // host-visible copy of the job's status
        status[active[blockIdx.x]] = BStatus{J.iter, J.state, J.shift};
    if (table)
        mfma::build_table_block<FP16>(J.cen[nxt], J.d, J.k, J.nt16, J.nkb, J.table, J.norm, J.stats, red);
    // FP32-scan jobs (state, shift and keep_ok are uniform: written before
    // the barrier).
    // Centroids that did not move keep their thresholds; a shift of 0 can
    // hide only a cluster that turned empty (finite -> NaN), whose old
    // centroid in the minimum only makes the other thresholds smaller.
    // MFMA jobs with d <= 32 also get their centroid rows every iteration.
    const bool fp32_keep = J.kind == 0 && (J.k >= kKeepMinK || J.k * J.d >= kKeepMinKd);
    const bool mfma_keep = J.kind == 1 && J.cen_t_w != 0 && (uint64_t)J.k * J.cen_t_w <= kKeepStage;
    if (J.state == CLK_JOB_ACTIVE && (fp32_keep || mfma_keep)) {
        const bool thr = !(J.keep_ok && J.shift == 0.0);
        if (thr || mfma_keep) {
            __syncthreads();
            build_keep_thresholds(J.cen[nxt], J.d, J.k, J.keep, cs, red, mfma_keep ? J.cen_t : nullptr, thr);
        }
        if (threadIdx.x == 0)
            J.keep_ok = 1;
    }
}

// Result copies of the serving loop: one work-group per (source, words,
// destination) triple; destinations in device-mapped pinned host memory.
struct GatherDesc {
    const uint32_t *src;
    uint32_t *dst;
    uint64_t words;
};

__global__ void __launch_bounds__(256) gather_kernel(const GatherDesc *__restrict__ descs)
{
    const GatherDesc g = descs[blockIdx.x];
    for (uint64_t i = threadIdx.x; i < g.words; i += blockDim.x)
        g.dst[i] = g.src[i];
}

// Activation of staged jobs (serving): one work-group per slot copies the
// job's record from device-mapped pinned host memory into the job table and,
// for the MFMA labeling, builds the first 16-bit table.
template <bool FP16>
__global__ void __launch_bounds__(256)
activate_kernel(const uint32_t *__restrict__ slots, const BJob *__restrict__ recs, BJob *__restrict__ jobs)
{
    __shared__ double red[256];
    const uint32_t s = slots[blockIdx.x];
    const uint32_t *src = reinterpret_cast<const uint32_t *>(recs + s);
    uint32_t *dst = reinterpret_cast<uint32_t *>(jobs + s);
    for (uint32_t i = threadIdx.x; i < sizeof(BJob) / 4; i += blockDim.x)
        dst[i] = src[i];
    const BJob &J = recs[s];
    if (J.kind == 1)
        mfma::build_table_block<FP16>(J.cen[0], J.d, J.k, J.nt16, J.nkb, J.table, J.norm, J.stats, red);
}

// The table of a job's current centroids (a newly admitted job: its initial
// ones; a job that iterated alone: those of its last iteration).
template <bool FP16>
__global__ void __launch_bounds__(256) batch_build_table_kernel(BJob *__restrict__ jobs, uint32_t slot)
{
    __shared__ double red[256];
    BJob &J = jobs[slot];
    mfma::build_table_block<FP16>(J.cen[J.iter & 1], J.d, J.k, J.nt16, J.nkb, J.table, J.norm, J.stats, red);
}

// The record update after an iteration of a job alone (the batch finish's
// bookkeeping for lloyd_step's result): shift, iteration count, stopping
// rule, status mirror.  The keep thresholds and centroid rows are now stale:
// the next batch finish recomputes them before any keep test reads them.
__global__ void __launch_bounds__(64)
solo_finish_kernel(BJob *__restrict__ jobs, uint32_t slot, const double *__restrict__ shift,
                   BStatus *__restrict__ status)
{
    if (threadIdx.x != 0)
        return;
    BJob &J = jobs[slot];
    J.shift = *shift;
    J.iter += 1;
    if (J.iter >= J.max_iter || (J.tol > 0.0 && J.shift < J.tol))
        J.state = CLK_JOB_DONE;
    J.keep_ok = 0;
    J.bound_ok = 0;  // the step wrote no bounds
    status[slot] = BStatus{J.iter, J.state, J.shift};
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
using clk::BMergeUnit;
using clk::BTile;

// A size class: the jobs of one tile-kernel instantiation.  Its LDS and
// work-group size follow the largest k and d among its active jobs.
struct BClassState {
    uint32_t tile0 = 0, ntiles = 0, grid = 0, waves = 0;
    size_t lds = 0;
    clk::BLayout layout{};
};

constexpr uint32_t kGraphAfter = 8;

struct clk_batch {
    clk_context *ctx = nullptr;
    clk_batch_config cfg{};
    bool mfma = false, fp16 = false;
    BJob *d_jobs = nullptr;
    BCtl *d_ctl = nullptr;  // one per class
    BTile *d_tiles = nullptr;
    uint32_t tile_cap = 0;
    uint32_t *d_active = nullptr;  // slots of the active jobs (finish work-groups)
    uint32_t nactive = 0;
    clk::BMergeUnit *d_units = nullptr;  // merge work-groups
    uint32_t nunits = 0, unit_cap = 0;
    uint32_t *d_retile = nullptr;  // (slot, tile_pts, ntiles) of jobs rebuild re-tiled
    uint32_t retile_cap = 0;
    char *arena = nullptr;
    // host mirror of the job table (device-side fields iter/state/shift are
    // read back by poll / take)
    std::vector<BJob> jobs;
    std::vector<uint32_t> staged_state;  // state to activate a staged slot with
    std::vector<uint64_t> off, bytes;  // arena ranges per slot
    uint32_t tile_work = 0;
    bool dirty = true;  // tile list / graph must be rebuilt
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    uint32_t stable_steps = 0;  // steps since the job set last changed
    bool dry = false;           // autotuning: iterations label, accumulate and merge only
    BClassState cls[clk::kClasses];
    // the size classes' tile kernels run concurrently: forked from the
    // step's stream onto these (default priority) and joined before the merge
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {};
    hipEvent_t fork = nullptr, join[kSide] = {};
    BJob *h_rec = nullptr;  // pinned staging of job records for activation
    BJob *h_rec_dev = nullptr;  // its device address
    // status of every slot after its last iteration, written by the finish
    // kernel (pinned, device-mapped): the serving loop reads it in place
    clk::BStatus *h_status = nullptr, *h_status_dev = nullptr;
};

namespace {

uint64_t job_bytes(uint64_t n, uint32_t d, uint32_t k, uint32_t tile_pts, bool points = true)
{
    const uint64_t nt = clk::div_up(std::max<uint64_t>(n, 1), tile_pts);
    const uint64_t nt16 = clk::div_up(k, clk::mfma::kTile), nkb = clk::div_up(d, clk::mfma::kKBlock);
    uint64_t b = 0;
    b += clk::align_up(nt16 * nkb * 64 * 16, 256);        // 16-bit table
    b += clk::align_up(nt16 * clk::mfma::kTile * 4, 256);  // norms
    b += 256;                                             // stats
    b += points ? clk::align_up(n * d * 4, 256) : 0;     // points (unless read in place)
    b += clk::align_up(n * 4, 256);                       // labels
    b += 2 * clk::align_up((uint64_t)d * k * 4, 256);     // centroids x2
    b += clk::align_up((uint64_t)k * 4, 256);             // masses
    b += clk::align_up((uint64_t)k * 4, 256);             // keep thresholds
    b += clk::align_up((uint64_t)k * clk::keep_row_width(d) * 4, 256);  // centroid rows
    b += clk::align_up(nt * (uint64_t)k * (d + 1) * 8, 256);  // partial rows
    b += clk::align_up((uint64_t)(d + 1) * 4, 256);            // grid (max |x| bits, NaN/inf seen)
    b += clk::align_up((uint64_t)d * k * 4, 256);              // off-grid flags
    b += d <= 8 ? clk::align_up(n * 4, 256) : 0;               // second-distance bounds
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

// The autotuner's tile_work candidates; a job's partial rows are allocated for
// the smallest, so the tuner can re-tile admitted jobs.
constexpr uint32_t kTileWorkCandidates[] = {4, 8, 16, 32, 64, 128, 256};
// rebuild halves a class's tiles while it has fewer tiles than this many per
// resident work-group
constexpr uint64_t kMinClassTiles = 4;
// tile_work when the configuration leaves it 0.  Large tiles write fewer
// partial rows; rebuild halves the tiles of a class that has too few of them
// for its work-groups (few jobs resident, as when serving), so a large
// default costs the sparse classes nothing.
constexpr uint32_t kDefaultTileWork = 128;

uint32_t row_tile_points(uint32_t tile_work, uint32_t d, uint32_t k)
{
    return tile_points(std::min(tile_work, kTileWorkCandidates[0]), d, k);
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

constexpr uint32_t kSmallK = 64;

// Host-side state of a slot whose copies are enqueued but which the kernels
// do not see yet (serving).
constexpr uint32_t kJobStaged = 3;

// Jobs with a small k*d label faster with the FP32 scan than on the matrix
// cores (the 16-bit path has a fixed cost per point: conversion, bounds,
// margin test, and d is padded to 32; the scan's keep test skips most of its
// work from the second iteration on).  Measured on the 4096-job mix (d in
// {3, 8, 16, 32}, k up to 256), solves/s by the k*d threshold: 1024: 18.6k,
// 1536: 19.4k, 2048: 20.4k, 3072: 21.3k, 4096: 20.5-21.2k, 6144: 19.7k,
// 12288: 17.1k.  With the exact sums, the second-distance bound and the
// d-equals-bucket tiles: 1536: 31.2k, 2048: 36.9-37.3k, 3072: 38.1k (a lower
// threshold shortens the first iteration, 10.5 -> 9.9 ms, but lengthens the
// later ones).  The labels are the same either way.
constexpr uint64_t kMfmaMinKD = 3072;

bool use_mfma(const clk_batch *b, uint32_t d, uint32_t k)
{
    return b->mfma && k >= 16 && (uint64_t)k * d >= kMfmaMinKD;
}

uint32_t class_of(uint32_t kind, uint32_t d, uint32_t k)
{
    const uint32_t big = k > kSmallK;
    if (kind == 0) {
        uint32_t i = 0;
        while (i + 1 < sizeof(clk::kFp32Dmax) / sizeof(clk::kFp32Dmax[0]) && clk::kFp32Dmax[i] < d)
            ++i;
        return 2 * i + big;
    }
    return clk::kMfmaClass0 + 2 * (d <= 16 ? 0u : d <= 32 ? 1u : 2u) + big;
}

const void *class_kernel(const clk_batch *b, uint32_t c)
{
    if (c < clk::kMfmaClass0) {
        switch (c / 2) {
        case 0: return (const void *)clk::batch_tiles_kernel<3>;
        case 1: return (const void *)clk::batch_tiles_kernel<4>;
        case 2: return (const void *)clk::batch_tiles_kernel<8>;
        case 3: return (const void *)clk::batch_tiles_kernel<16>;
        case 4: return (const void *)clk::batch_tiles_kernel<32>;
        default: return (const void *)clk::batch_tiles_kernel<64>;
        }
    }
    const bool nkb2 = c >= clk::kMfmaNkb2Class;
    if (b->fp16)
        return nkb2 ? (const void *)clk::batch_tiles_mfma_kernel<true, 2>
                    : (const void *)clk::batch_tiles_mfma_kernel<true, 1>;
    return nkb2 ? (const void *)clk::batch_tiles_mfma_kernel<false, 2>
                : (const void *)clk::batch_tiles_mfma_kernel<false, 1>;
}

// LDS of class c for jobs up to kmax clusters and dmax features with the
// given waves per work-group.
size_t class_lds(uint32_t c, uint32_t waves, uint32_t kmax, uint32_t kdmax, uint32_t dmax, clk::BLayout *layout)
{
    if (c < clk::kMfmaClass0)
        return clk::batch_lds_bytes(waves, kmax, kdmax, clk::kFp32Dmax[c / 2]);
    const clk::BLayout L = clk::mfma_layout_for(waves, kmax, kdmax, dmax);
    if (layout)
        *layout = L;
    return L.bytes;
}

// The most waves per work-group whose LDS fits (at most 4 for the FP32
// scan, 8 for the one-K-block MFMA classes, whose work-group shares one
// accumulator and the 16-bit table); 0 if none does.
uint32_t class_waves(const clk_batch *b, uint32_t c, uint32_t kmax, uint32_t kdmax, uint32_t dmax)
{
    // (two K-blocks need more than 256 VGPRs: four waves)
    const uint32_t most = c < clk::kMfmaClass0 || c >= clk::kMfmaNkb2Class ? clk::kBatchMaxWaves
                                                                             : clk::kBatchMfmaWaves;
    for (uint32_t w = most; w >= 1; --w)
        if (class_lds(c, w, kmax, kdmax, dmax, nullptr) + 4096 <= b->ctx->lds_per_block)  // + static LDS
            return w;
    return 0;
}

// Whether class c still fits with one more job (k, d) beside its active and
// staged jobs (the sums of a tile need k*d int64 cells in LDS).
bool class_fits_with(const clk_batch *b, uint32_t c, uint32_t k, uint32_t d)
{
    uint32_t kmax = k, kdmax = k * d, dmax = d;
    for (size_t s = 0; s < b->jobs.size(); ++s) {
        const BJob &J = b->jobs[s];
        if ((J.state != CLK_JOB_ACTIVE && J.state != kJobStaged) || class_of(J.kind, J.d, J.k) != c)
            continue;
        kmax = std::max(kmax, J.k);
        kdmax = std::max(kdmax, J.k * J.d);
        dmax = std::max(dmax, J.d);
    }
    return class_waves(b, c, kmax, kdmax, dmax) > 0;
}

int rebuild(clk_batch *b, hipStream_t stream)
{
    std::vector<BTile> tiles;
    std::vector<uint32_t> active, retile;
    std::vector<BCtl> ctl(clk::kClasses, BCtl{0, 0, b->dry ? 1u : 0u, 0});
    for (uint32_t c = 0; c < clk::kClasses; ++c) {
        BClassState &C = b->cls[c];
        C = BClassState{};
        C.tile0 = (uint32_t)tiles.size();
        std::vector<uint32_t> mem;
        uint32_t kmax = 0, kdmax = 0, dmax = 0;
        for (uint32_t s = 0; s < b->jobs.size(); ++s) {
            const BJob &J = b->jobs[s];
            if (J.state != CLK_JOB_ACTIVE || class_of(J.kind, J.d, J.k) != c)
                continue;
            mem.push_back(s);
            kmax = std::max(kmax, J.k);
            kdmax = std::max(kdmax, J.k * J.d);
            dmax = std::max(dmax, J.d);
        }
        if (mem.empty())
            continue;
        // the most waves whose LDS fits the class's jobs (the sums are
        // exact, so the wave count does not change any result)
        C.waves = class_waves(b, c, kmax, kdmax, dmax);
        if (C.waves == 0) {  // admission checks every job's class (class_fits_with)
            clk::set_error("clk_batch: a size class does not fit the work-group's LDS");
            return CLK_ERR_UNSUPPORTED;
        }
        C.lds = class_lds(c, C.waves, kmax, kdmax, dmax, &C.layout);
        const void *kern = class_kernel(b, c);
        if (C.lds > 65536)
            CLK_HIP_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C.lds));
        // Tile sizes: each job's base size (the batch's tile_work), halved
        // -- not below the smallest candidate's, which its partial rows are
        // sized for -- while the class has fewer than kMinClassTiles tiles
        // per resident work-group: a class of few, large tiles leaves work-
        // groups idle behind its last ones.  (The sums are exact: the tiling
        // changes no result.)
        const uint64_t resident = clk::auto_grid(b->ctx, kern, (int)(64 * C.waves), C.lds, ~0ull >> 8);
        std::vector<uint32_t> tp(mem.size());
        uint64_t total = 0;
        for (size_t i = 0; i < mem.size(); ++i) {
            const BJob &J = b->jobs[mem[i]];
            tp[i] = J.tile_base;
            total += clk::div_up(J.n, (uint64_t)tp[i]);
        }
        while (total < kMinClassTiles * resident) {
            bool shrunk = false;
            total = 0;
            for (size_t i = 0; i < mem.size(); ++i) {
                const BJob &J = b->jobs[mem[i]];
                const uint32_t floor_pts = tile_points(kTileWorkCandidates[0], J.d, J.k);
                const uint32_t half = std::max(floor_pts, tp[i] / 2 / 256 * 256);
                if (half < tp[i]) {
                    tp[i] = half;
                    shrunk = true;
                }
                total += clk::div_up(J.n, (uint64_t)tp[i]);
            }
            if (!shrunk)
                break;
        }
        for (size_t i = 0; i < mem.size(); ++i) {
            const uint32_t s = mem[i];
            BJob &J = b->jobs[s];
            if (J.tile_pts != tp[i]) {
                J.tile_pts = tp[i];
                J.ntiles = (uint32_t)clk::div_up(J.n, (uint64_t)tp[i]);
                retile.insert(retile.end(), {s, J.tile_pts, J.ntiles});
            }
            active.push_back(s);
            J.tile0 = (uint32_t)tiles.size();
            for (uint32_t t = 0; t < J.ntiles; ++t)
                tiles.push_back({s, t});
        }
        C.ntiles = (uint32_t)tiles.size() - C.tile0;
        ctl[c].ntiles = C.ntiles;
        C.grid = (uint32_t)clk::auto_grid(b->ctx, kern, (int)(64 * C.waves), C.lds,
                                          (uint64_t)C.ntiles * 64 * C.waves);
    }
    if (!retile.empty()) {  // the device records' tiling
        const uint32_t count = (uint32_t)(retile.size() / 3);
        if (retile.size() > b->retile_cap) {
            if (b->d_retile)
                CLK_HIP_CHECK(hipFree(b->d_retile));
            b->retile_cap = (uint32_t)std::max<size_t>(retile.size(), 3 * 1024) * 2;
            CLK_HIP_CHECK(hipMalloc(&b->d_retile, b->retile_cap * sizeof(uint32_t)));
        }
        CLK_HIP_CHECK(hipMemcpyAsync(b->d_retile, retile.data(), retile.size() * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(clk::batch_retile_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, b->d_jobs,
                           b->d_retile, count);
        CLK_HIP_CHECK(hipGetLastError());
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
    if (!active.empty())
        CLK_HIP_CHECK(hipMemcpyAsync(b->d_active, active.data(), active.size() * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, stream));
    b->nactive = (uint32_t)active.size();
    std::vector<BMergeUnit> units;
    for (uint32_t s : active) {
        const BJob &J = b->jobs[s];
        // a job of one tile has its sums in row 0 already (its values are
        // counts and unit multiples, never -0, so 0 + row 0 would change no
        // bit)
        if (J.ntiles <= 1)
            continue;
        const uint32_t rl = J.k * (J.d + 1);
        const uint32_t per = clk::kMergeThreads / clk::merge_split(rl, J.ntiles);
        for (uint32_t e0 = 0; e0 < rl; e0 += per)
            units.push_back({s, e0});
    }
    if (units.size() > b->unit_cap) {
        if (b->d_units)
            CLK_HIP_CHECK(hipFree(b->d_units));
        b->unit_cap = (uint32_t)std::max<size_t>(units.size(), 1024) * 2;
        CLK_HIP_CHECK(hipMalloc(&b->d_units, b->unit_cap * sizeof(BMergeUnit)));
    }
    if (!units.empty())
        CLK_HIP_CHECK(hipMemcpyAsync(b->d_units, units.data(), units.size() * sizeof(BMergeUnit),
                                     hipMemcpyHostToDevice, stream));
    b->nunits = (uint32_t)units.size();
    CLK_HIP_CHECK(hipMemcpyAsync(b->d_ctl, ctl.data(), ctl.size() * sizeof(BCtl), hipMemcpyHostToDevice, stream));
    CLK_HIP_CHECK(hipStreamSynchronize(stream));  // tiles / ctl are host temporaries
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

// The tile kernel of every non-empty class, then the merge and finish
// kernels (the finish resets the tile counters).  The classes are
// independent (their own jobs, tile counters and partial rows): they run
// concurrently on the step's stream and the batch's side streams, forked
// and joined by events (inside a graph capture these become its edges), so
// the tail of one class overlaps the others.
int launch_step(clk_batch *b, hipStream_t stream_main)
{
    uint32_t used = 0;  // side streams forked so far
    uint32_t nlaunched = 0;
    for (uint32_t c = 0; c < clk::kClasses; ++c) {
        const BClassState &C = b->cls[c];
        if (C.ntiles == 0)
            continue;
        // round robin: main stream, then the side streams
        const uint32_t slot = nlaunched++ % (clk_batch::kSide + 1);
        hipStream_t stream = stream_main;
        if (slot != 0) {
            const uint32_t si = slot - 1;
            if (si >= used) {
                if (used == 0)
                    CLK_HIP_CHECK(hipEventRecord(b->fork, stream_main));
                CLK_HIP_CHECK(hipStreamWaitEvent(b->side[si], b->fork, 0));
                used = si + 1;
            }
            stream = b->side[si];
        }
        const BTile *tl = b->d_tiles + C.tile0;
        BCtl *ctl = b->d_ctl + c;
        const dim3 g(C.grid), t(64 * C.waves);
        if (c < clk::kMfmaClass0) {
            switch (c / 2) {
            case 0: hipLaunchKernelGGL(clk::batch_tiles_kernel<3>, g, t, C.lds, stream, b->d_jobs, tl, ctl); break;
            case 1: hipLaunchKernelGGL(clk::batch_tiles_kernel<4>, g, t, C.lds, stream, b->d_jobs, tl, ctl); break;
            case 2: hipLaunchKernelGGL(clk::batch_tiles_kernel<8>, g, t, C.lds, stream, b->d_jobs, tl, ctl); break;
            case 3: hipLaunchKernelGGL(clk::batch_tiles_kernel<16>, g, t, C.lds, stream, b->d_jobs, tl, ctl); break;
            case 4: hipLaunchKernelGGL(clk::batch_tiles_kernel<32>, g, t, C.lds, stream, b->d_jobs, tl, ctl); break;
            default: hipLaunchKernelGGL(clk::batch_tiles_kernel<64>, g, t, C.lds, stream, b->d_jobs, tl, ctl); break;
            }
        } else if (b->fp16) {
            if (c >= clk::kMfmaNkb2Class)
                hipLaunchKernelGGL((clk::batch_tiles_mfma_kernel<true, 2>), g, t, C.lds, stream, b->d_jobs, tl,
                                   ctl, C.layout);
            else
                hipLaunchKernelGGL((clk::batch_tiles_mfma_kernel<true, 1>), g, t, C.lds, stream, b->d_jobs, tl,
                                   ctl, C.layout);
        } else {
            if (c >= clk::kMfmaNkb2Class)
                hipLaunchKernelGGL((clk::batch_tiles_mfma_kernel<false, 2>), g, t, C.lds, stream, b->d_jobs, tl,
                                   ctl, C.layout);
            else
                hipLaunchKernelGGL((clk::batch_tiles_mfma_kernel<false, 1>), g, t, C.lds, stream, b->d_jobs, tl,
                                   ctl, C.layout);
        }
        CLK_HIP_CHECK(hipGetLastError());
    }
    for (uint32_t si = 0; si < used; ++si) {
        CLK_HIP_CHECK(hipEventRecord(b->join[si], b->side[si]));
        CLK_HIP_CHECK(hipStreamWaitEvent(stream_main, b->join[si], 0));
    }
    hipStream_t stream = stream_main;
    if (b->nunits > 0) {
        hipLaunchKernelGGL(clk::batch_merge_kernel, dim3(b->nunits), dim3(clk::kMergeThreads), 0, stream, b->d_jobs,
                           b->d_units, b->d_ctl);
        CLK_HIP_CHECK(hipGetLastError());
    }
    const dim3 fg(std::max<uint32_t>(b->nactive, 1)), fb(256);
    if (b->fp16)
        hipLaunchKernelGGL(clk::batch_finish_kernel<true>, fg, fb, 0, stream, b->d_jobs, b->d_active, b->d_ctl,
                           b->h_status_dev);
    else
        hipLaunchKernelGGL(clk::batch_finish_kernel<false>, fg, fb, 0, stream, b->d_jobs, b->d_active, b->d_ctl,
                           b->h_status_dev);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

}  // namespace

extern "C" uint64_t clk_batch_job_bytes(const clk_batch_config *cfg, uint64_t n, uint32_t d, uint32_t k)
{
    const uint32_t tw = (cfg && cfg->tile_work) ? cfg->tile_work : kDefaultTileWork;
    return job_bytes(n, d, k, row_tile_points(tw, d, k));
}

extern "C" int clk_batch_create(clk_context *ctx, const clk_batch_config *cfg, clk_batch **out)
{
    CLK_TRY(clk::enter(ctx));
    CLK_REQUIRE(cfg && out, CLK_ERR_INVALID_ARGUMENT, "clk_batch_create: null argument");
    CLK_REQUIRE(cfg->max_jobs >= 1 && cfg->arena_bytes > 0, CLK_ERR_SIZE, "clk_batch_create: empty batch");
    CLK_REQUIRE(cfg->max_d >= 1 && cfg->max_d <= 64 && cfg->max_k >= 1 && cfg->max_k <= (1u << 20),
                CLK_ERR_SIZE, "clk_batch_create: max_d must be 1..64, max_k 1..2^20");
    *out = nullptr;
    clk_batch *b = new clk_batch;
    b->ctx = ctx;
    b->cfg = *cfg;
    b->tile_work = cfg->tile_work ? cfg->tile_work : kDefaultTileWork;
    CLK_REQUIRE(cfg->labeling == CLK_LABELING_UNROLL_VECTOR || cfg->labeling == CLK_LABELING_MFMA_BF16 ||
                    cfg->labeling == CLK_LABELING_MFMA_FP16,
                CLK_ERR_INVALID_ARGUMENT, "clk_batch_create: unknown labeling strategy id");
    b->mfma = cfg->labeling != CLK_LABELING_UNROLL_VECTOR;
    b->fp16 = cfg->labeling == CLK_LABELING_MFMA_FP16;
    // every class must take jobs up to 64 clusters with max_d features (a
    // job with more clusters is checked at admission: a tile's exact sums
    // need k*d int64 cells in LDS)
    {
        bool fits = cfg->max_d >= 1 && cfg->max_d <= 64 && cfg->max_k >= 1;
        for (uint32_t kind = 0; fits && kind <= (b->mfma ? 1u : 0u); ++kind) {
            const uint32_t k = std::min(cfg->max_k, kSmallK);
            const uint32_t c = class_of(kind, cfg->max_d, k);
            fits = class_waves(b, c, k, k * cfg->max_d, cfg->max_d) > 0;
        }
        if (!fits) {
            delete b;
            clk::set_error("clk_batch_create: max_d outside 1..64, or a size class does not fit the LDS");
            return CLK_ERR_UNSUPPORTED;
        }
    }
    if (b->mfma) {
        const int rc = clk::mfma_subnormal_flags(ctx, nullptr);
        if (rc != CLK_OK) {
            delete b;
            return rc;
        }
    }
    b->jobs.assign(cfg->max_jobs, BJob{});
    b->staged_state.assign(cfg->max_jobs, 0);
    b->off.assign(cfg->max_jobs, 0);
    b->bytes.assign(cfg->max_jobs, 0);
    hipError_t e1 = hipMalloc(&b->d_jobs, cfg->max_jobs * sizeof(BJob));
    hipError_t e2 = hipMalloc(&b->d_ctl, clk::kClasses * sizeof(BCtl));
    hipError_t e3 = hipMalloc(&b->arena, cfg->arena_bytes);
    hipError_t e4 = hipMalloc(&b->d_active, cfg->max_jobs * sizeof(uint32_t));
    hipError_t e5 = hipHostMalloc((void **)&b->h_status, cfg->max_jobs * sizeof(clk::BStatus), hipHostMallocDefault);
    if (e5 == hipSuccess)
        e5 = hipHostGetDevicePointer((void **)&b->h_status_dev, b->h_status, 0);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess || e5 != hipSuccess) {
        clk_batch_destroy(b);
        clk::set_error("clk_batch_create: device allocation failed");
        return CLK_ERR_NOMEM;
    }
    CLK_HIP_CHECK(hipMemset(b->d_jobs, 0, cfg->max_jobs * sizeof(BJob)));
    CLK_HIP_CHECK(hipMemset(b->d_ctl, 0, clk::kClasses * sizeof(BCtl)));
    bool ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < clk_batch::kSide && ok; ++i)
        ok = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        clk_batch_destroy(b);
        clk::set_error("clk_batch_create: stream / event creation failed");
        return CLK_ERR_HIP;
    }
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
    if (b->d_active)
        (void)hipFree(b->d_active);
    if (b->d_units)
        (void)hipFree(b->d_units);
    if (b->d_retile)
        (void)hipFree(b->d_retile);
    if (b->h_rec)
        (void)hipHostFree(b->h_rec);
    if (b->h_status)
        (void)hipHostFree(b->h_status);
    if (b->arena)
        (void)hipFree(b->arena);
    for (int i = 0; i < clk_batch::kSide; ++i) {
        if (b->side[i])
            (void)hipStreamDestroy(b->side[i]);
        if (b->join[i])
            (void)hipEventDestroy(b->join[i]);
    }
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    delete b;
    return CLK_OK;
}

namespace {

// Host-side slot state while a job's data is copied in but the job is not
// yet in the device job table (the serving loop stages copies while the GPU
// works).

// Stage a job: a free slot and arena range, the job record, and the copies of
// its points and initial centroids (async on stream).  The slot is not yet
// visible to the kernels.
int stage_job(clk_batch *b, uint64_t n, uint32_t d, uint32_t k, const float *host_points,
              const float *host_centroids, uint32_t max_iterations, double tolerance, int32_t *slot,
              hipStream_t st, bool device_points = false, bool zero_outputs = true)
{
    CLK_REQUIRE(n >= 1 && n <= (1ull << 30) && n * d <= (1ull << 30), CLK_ERR_SIZE,
                "clk_batch_add: n must be >= 1 with n*d <= 2^30 (32-bit byte offsets into the points)");
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
    {
        const uint32_t cls = class_of(use_mfma(b, d, k) ? 1u : 0u, d, k);
        if (!class_fits_with(b, cls, k, d)) {
            // alone it never fits; beside the resident jobs of its class it
            // fits once some of them have finished
            const bool alone = class_waves(b, cls, k, k * d, d) > 0;
            clk::set_error(alone ? "clk_batch_add: the job's size class is full (LDS); retry after jobs finish"
                                 : "clk_batch_add: k*d too large for the batch engine's LDS accumulator");
            return alone ? CLK_ERR_NOMEM : CLK_ERR_UNSUPPORTED;
        }
    }
    const uint32_t tp = tile_points(b->tile_work, d, k);
    const uint64_t need = job_bytes(n, d, k, row_tile_points(b->tile_work, d, k), !device_points);
    uint64_t at = 0;
    if (!arena_alloc(b, need, &at)) {
        clk::set_error("clk_batch_add: arena full");
        return CLK_ERR_NOMEM;
    }
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
    if (device_points) {
        J.pts = host_points;
    } else {
        J.pts = reinterpret_cast<const float *>(p);
        p += clk::align_up(n * d * 4, 256);
    }
    J.labels = reinterpret_cast<uint32_t *>(p);
    p += clk::align_up(n * 4, 256);
    J.cen[0] = reinterpret_cast<float *>(p);
    p += clk::align_up((uint64_t)d * k * 4, 256);
    J.cen[1] = reinterpret_cast<float *>(p);
    p += clk::align_up((uint64_t)d * k * 4, 256);
    J.masses = reinterpret_cast<uint32_t *>(p);
    p += clk::align_up((uint64_t)k * 4, 256);
    J.keep = reinterpret_cast<float *>(p);
    p += clk::align_up((uint64_t)k * 4, 256);
    J.keep_ok = 0;  // the first iteration scans every centroid
    J.cen_t = reinterpret_cast<float *>(p);
    J.cen_t_w = clk::keep_row_width(d);
    p += clk::align_up((uint64_t)k * J.cen_t_w * 4, 256);
    J.part = reinterpret_cast<double *>(p);
    p += clk::align_up(clk::div_up(n, row_tile_points(b->tile_work, d, k)) * (uint64_t)k * (d + 1) * 8, 256);
    J.gmax = reinterpret_cast<uint32_t *>(p);
    p += clk::align_up((uint64_t)(d + 1) * 4, 256);
    J.flags = reinterpret_cast<uint32_t *>(p);
    p += clk::align_up((uint64_t)d * k * 4, 256);
    // the second-distance bounds of the d <= 8 FP32 tiles (reserved for every
    // d <= 8 job, used by the FP32-scan ones)
    J.lb = d <= 8 && !use_mfma(b, d, k) ? reinterpret_cast<float *>(p) : nullptr;
    J.bound_ok = 0;  // written by the first iteration's tiles
    J.drift = 0.0f;
    J.n = n;
    J.d = d;
    J.k = k;
    J.tile_pts = tp;
    J.tile_base = tp;
    J.ntiles = (uint32_t)clk::div_up(n, tp);
    J.max_iter = max_iterations;
    J.iter = 0;
    J.state = max_iterations > 0 ? CLK_JOB_ACTIVE : CLK_JOB_DONE;
    J.dmax = clk::bucket_of(d);
    J.kind = use_mfma(b, d, k) ? 1u : 0u;
    J.tol = tolerance;
    J.shift = 0.0;
    if (!device_points)
        CLK_HIP_CHECK(
            hipMemcpyAsync(const_cast<float *>(J.pts), host_points, n * d * 4, hipMemcpyHostToDevice, st));
    CLK_HIP_CHECK(hipMemcpyAsync(J.cen[0], host_centroids, (size_t)d * k * 4, hipMemcpyHostToDevice, st));
    // the job's sum grid (the one clk_sum_grid_from_points computes for these
    // points) and cleared flags
    CLK_TRY(clk::launch_grid_max(b->ctx, d, n, J.pts, J.gmax, st));
    CLK_HIP_CHECK(hipMemsetAsync(J.flags, 0, (size_t)d * k * 4, st));
    if (zero_outputs || max_iterations == 0) {  // else the first iteration writes both
        CLK_HIP_CHECK(hipMemsetAsync(J.masses, 0, (size_t)k * 4, st));
        CLK_HIP_CHECK(hipMemsetAsync(J.labels, 0, n * 4, st));
    }
    b->jobs[s] = J;
    b->jobs[s].state = kJobStaged;
    b->staged_state[s] = J.state;
    b->off[s] = at;
    b->bytes[s] = need;
    *slot = s;
    return CLK_OK;
}

// Pinned, device-mapped copies of the job records (the copies and the
// activation kernel read them after the call that wrote them returns).
int ensure_pinned_records(clk_batch *b)
{
    if (b->h_rec)
        return CLK_OK;
    CLK_HIP_CHECK(hipHostMalloc((void **)&b->h_rec, b->jobs.size() * sizeof(BJob), hipHostMallocDefault));
    CLK_HIP_CHECK(hipHostGetDevicePointer((void **)&b->h_rec_dev, b->h_rec, 0));
    return CLK_OK;
}

// Make a staged job visible: its record in the device job table and, for the
// MFMA labeling, its first 16-bit table (stream-ordered after the copies when
// st is the staging stream; otherwise the caller orders the two streams).
// No step may be in flight on another stream.
int activate_job(clk_batch *b, int32_t s, hipStream_t st, bool sync = true)
{
    BJob &J = b->jobs[s];
    J.state = b->staged_state[s];
    CLK_TRY(ensure_pinned_records(b));
    b->h_rec[s] = J;
    CLK_HIP_CHECK(hipMemcpyAsync(b->d_jobs + s, b->h_rec + s, sizeof(BJob), hipMemcpyHostToDevice, st));
    if (J.kind == 1) {
        if (b->fp16)
            hipLaunchKernelGGL(clk::batch_build_table_kernel<true>, dim3(1), dim3(256), 0, st, b->d_jobs,
                               (uint32_t)s);
        else
            hipLaunchKernelGGL(clk::batch_build_table_kernel<false>, dim3(1), dim3(256), 0, st, b->d_jobs,
                               (uint32_t)s);
        CLK_HIP_CHECK(hipGetLastError());
    }
    // the pinned record is not rewritten before the slot is taken again,
    // which synchronizes the stream first
    if (sync)
        CLK_HIP_CHECK(hipStreamSynchronize(st));
    b->dirty = true;
    return CLK_OK;
}

}  // namespace

extern "C" int clk_batch_add(clk_batch *b, uint64_t n, uint32_t d, uint32_t k, const float *host_points,
                             const float *host_centroids, uint32_t max_iterations, double tolerance,
                             int32_t *slot, void *stream)
{
    CLK_REQUIRE(b && slot, CLK_ERR_INVALID_ARGUMENT, "clk_batch_add: null argument");
    CLK_TRY(clk::enter(b->ctx));
    hipStream_t st = clk::as_stream(stream);
    int32_t s = -1;
    CLK_TRY(stage_job(b, n, d, k, host_points, host_centroids, max_iterations, tolerance, &s, st));
    const int rc = activate_job(b, s, st);  // synchronizes: host_points may be reused at once
    if (rc != CLK_OK) {
        b->jobs[s].state = CLK_JOB_FREE;
        return rc;
    }
    *slot = s;
    return CLK_OK;
}

extern "C" int clk_batch_step(clk_batch *b, void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_step: null batch");
    CLK_TRY(clk::enter(b->ctx));
    hipStream_t st = clk::as_stream(stream);
    if (b->dirty) {
        CLK_TRY(rebuild(b, st));
        b->stable_steps = 0;
    }
    if (b->nactive == 0)
        return CLK_OK;
    // a graph pays off once the job set repeats: while jobs come and go
    // (serving), launch directly.  Capture + instantiation cost more than a
    // few direct launches, so a graph is made after kGraphAfter steps with
    // the same jobs (serving at 300 jobs/s: p50 0.58 ms capturing at the
    // second step, 0.43 ms at the eighth)
    if (++b->stable_steps < kGraphAfter) {
        clk::KernelTimer t(b->ctx, "BatchStep", st);
        return launch_step(b, st);
    }
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
        // the job's sum grid (as the batch's admission computes it); the
        // job's partial rows double as its FP64 sums buffer [d][k]
        rc = clk_sum_grid_from_points(b->ctx, J.d, J.n, J.pts, st);
        if (rc != CLK_OK)
            break;
        while (J.state == CLK_JOB_ACTIVE) {
            float *cur = J.cen[J.iter & 1], *nxt = J.cen[(J.iter & 1) ^ 1];
            rc = clk_lloyd_step(b->ctx, &cfg, J.d, J.n, J.k, J.pts, cur, nxt, J.part, J.labels,
                                J.masses, J.tol > 0.0 ? shift_d : nullptr, st);
            if (rc != CLK_OK)
                break;
            J.iter += 1;
            J.keep_ok = 0;  // the keep thresholds are of older centroids
            J.bound_ok = 0;  // and the step wrote no bounds
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
    (void)clk_sum_grid_clear(b->ctx);
    CLK_TRY(rc);
    CLK_HIP_CHECK(hipMemcpy(b->d_jobs, dev.data(), dev.size() * sizeof(BJob), hipMemcpyHostToDevice));
    return CLK_OK;
}

// ---------------------------------------------------------------- serving ---

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void latency_stats(const clk_serve_job *jobs, uint32_t count, clk_serve_stats *st)
{
    std::vector<double> lat(count);
    double first = jobs[0].arrival, last = 0.0, sum = 0.0;
    for (uint32_t i = 0; i < count; ++i) {
        lat[i] = jobs[i].finished - jobs[i].arrival;
        sum += lat[i];
        first = std::min(first, jobs[i].arrival);
        last = std::max(last, jobs[i].finished);
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {  // nearest rank
        const size_t r = (size_t)std::ceil(q * (double)count);
        return lat[std::min<size_t>(count - 1, r > 0 ? r - 1 : 0)];
    };
    st->makespan = last - first;
    st->solves_per_s = st->makespan > 0.0 ? (double)count / st->makespan : 0.0;
    st->p50 = pct(0.50);
    st->p99 = pct(0.99);
    st->mean_latency = sum / (double)count;
}

int check_serve_jobs(const clk_serve_job *jobs, uint32_t count)
{
    CLK_REQUIRE(jobs && count > 0, CLK_ERR_INVALID_ARGUMENT, "clk_serve: no jobs");
    for (uint32_t i = 0; i < count; ++i) {
        CLK_REQUIRE(jobs[i].points && jobs[i].centroids, CLK_ERR_INVALID_ARGUMENT, "clk_serve: null job data");
        CLK_REQUIRE(i == 0 || jobs[i].arrival >= jobs[i - 1].arrival, CLK_ERR_INVALID_ARGUMENT,
                    "clk_serve: jobs must be in arrival order");
    }
    return CLK_OK;
}

struct Staged {
    int32_t slot;
    uint32_t job;
    hipEvent_t copied;
};

}  // namespace

extern "C" int clk_serve_run(clk_batch *b, clk_serve_job *jobs, uint32_t count, clk_serve_stats *stats,
                             void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_serve_run: null batch");
    CLK_TRY(clk::enter(b->ctx));
    CLK_TRY(check_serve_jobs(jobs, count));
    hipStream_t st = clk::as_stream(stream);
    CLK_REQUIRE(st != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_serve_run: needs a created stream");
    for (uint32_t s = 0; s < b->jobs.size(); ++s)
        CLK_REQUIRE(b->jobs[s].state == CLK_JOB_FREE, CLK_ERR_INVALID_ARGUMENT, "clk_serve_run: batch not empty");
    for (uint32_t i = 0; i < count; ++i) {  // every job fits an empty batch, so admission cannot stall
        const clk_serve_job &j = jobs[i];
        CLK_REQUIRE(j.n >= 1 && j.d >= 1 && j.d <= b->cfg.max_d && j.k >= 1 && j.k <= b->cfg.max_k &&
                        j.n <= (1ull << 30) && j.n * j.d <= (1ull << 30),
                    CLK_ERR_SIZE,
                    "clk_serve_run: a job's n, d or k is outside the batch's limits");
        CLK_REQUIRE(job_bytes(j.n, j.d, j.k, row_tile_points(b->tile_work, j.d, j.k), j.device_points == 0) <=
                        b->cfg.arena_bytes,
                    CLK_ERR_NOMEM, "clk_serve_run: a job is larger than the arena");
        CLK_REQUIRE(class_waves(b, class_of(use_mfma(b, j.d, j.k) ? 1u : 0u, j.d, j.k), j.k, j.k * j.d, j.d) > 0,
                    CLK_ERR_UNSUPPORTED, "clk_serve_run: a job's k*d is too large for the batch engine's LDS");
    }
    const uint32_t slots = (uint32_t)b->jobs.size();
    hipStream_t cs = nullptr;  // admission copies, concurrent with the iterations
    hipStream_t rs = nullptr;  // result copies, concurrent with the next iteration
    hipEvent_t step_done = nullptr;
    std::vector<clk::BStatus> done_status(slots);
    char *res = nullptr, *res_dev = nullptr;  // pinned, device-mapped result staging
    clk::GatherDesc *descs = nullptr, *descs_dev = nullptr;
    size_t res_cap = 0, desc_cap = 0;
    std::vector<uint32_t> done_slots, act_slots;
    std::vector<Staged> zero_iter;
    uint32_t *h_act = nullptr, *h_act_dev = nullptr;  // pinned slot list of the activation kernel
    size_t act_cap = 0;
    CLK_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    if (hipStreamCreateWithFlags(&rs, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        clk::set_error("clk_serve_run: stream creation failed");
        return CLK_ERR_HIP;
    }
    int rc = CLK_OK;
    std::vector<int64_t> slot_job(slots, -1);
    // A lone small job (nothing else resident, staged or arrived) iterates
    // through the plain per-job fused pass on its own buffers and grid: the
    // same results bit for bit (exact labels, exact sums on the job's grid),
    // one streaming launch plus its merge instead of the batch's tile, merge
    // and finish launches, whose few long tiles leave most of the GPU idle
    // at low load.  host_iter tracks each job's iteration count (which
    // centroid buffer is current); solo_ran marks jobs whose MFMA table and
    // keep data must be rebuilt before they rejoin a batch step.
    constexpr uint64_t kSoloMaxPoints = 1ull << 20;
    clk_kernel_config solo_cfg{};
    double *solo_shift = nullptr;
    std::vector<uint32_t> host_iter(slots, 0);
    std::vector<uint8_t> in_step(slots, 0), solo_ran(slots, 0);
    std::deque<Staged> staged;
    auto fail = [&](int code) {
        rc = code;
        return code;
    };
    if (hipEventCreateWithFlags(&step_done, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&solo_shift, sizeof(double)) != hipSuccess) {
        clk::set_error("clk_serve_run: event / device allocation failed");
        fail(CLK_ERR_HIP);
    }
    if (rc == CLK_OK && clk_default_config(b->ctx, CLK_STAGE_FUSED, CLK_FUSED_CLUSTER_MERGE, &solo_cfg) != CLK_OK)
        fail(CLK_ERR_INVALID_ARGUMENT);
    if (rc == CLK_OK && ensure_pinned_records(b) != CLK_OK)
        fail(CLK_ERR_HIP);
    const auto t0 = std::chrono::steady_clock::now();
    uint32_t next = 0, finished = 0, active = 0;
    bool inflight = false;
    uint64_t steps = 0;
    uint32_t max_active = 0;
    uint64_t solo_steps = 0;
    while (rc == CLK_OK && finished < count) {
        const double now = seconds_since(t0);
        // 1. stage the arrived jobs in arrival order while slots and arena allow
        //    (their copies overlap the iteration in flight)
        const size_t round0 = staged.size();
        while (next < count && jobs[next].arrival <= now) {
            clk_serve_job &j = jobs[next];
            int32_t s = -1;
            const int r = stage_job(b, j.n, j.d, j.k, j.points, j.centroids, j.max_iterations, j.tolerance, &s, cs,
                                    j.device_points != 0, false);
            if (r == CLK_ERR_NOMEM)
                break;
            if (r != CLK_OK) {
                fail(r);
                break;
            }
            j.admitted = seconds_since(t0);
            slot_job[s] = next;
            staged.push_back({s, next, nullptr});
            ++next;
        }
        if (rc != CLK_OK)
            break;
        if (staged.size() > round0) {  // one event marks this round's copies
            hipEvent_t ev = nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(ev, cs) != hipSuccess) {
                clk::set_error("clk_serve_run: event record failed");
                fail(CLK_ERR_HIP);
                break;
            }
            for (size_t i = round0; i < staged.size(); ++i)
                staged[i].copied = ev;
        }
        // 2. end of the iteration in flight: collect finished jobs
        if (inflight) {
            if (hipEventSynchronize(step_done) != hipSuccess) {
                clk::set_error("clk_serve_run: iteration failed");
                fail(CLK_ERR_HIP);
                break;
            }
            inflight = false;
            // the finish kernel of that iteration wrote every active job's
            // status to b->h_status (pinned): read in place.  Finished jobs
            // leave the tile list now; their results are copied while the
            // next iteration runs (no kernel touches a DONE job)
            const clk::BStatus *h_status = b->h_status;
            done_slots.clear();
            for (uint32_t s = 0; s < slots; ++s)
                if (in_step[s])
                    host_iter[s] = h_status[s].iter;
            for (uint32_t s = 0; s < slots; ++s)
                if (b->jobs[s].state == CLK_JOB_ACTIVE && h_status[s].state == CLK_JOB_DONE) {
                    b->jobs[s].state = CLK_JOB_DONE;
                    b->dirty = true;
                    done_slots.push_back(s);
                    done_status[s] = h_status[s];
                    --active;
                }
        }
        // 3. activate staged jobs whose copies are done (all of them when the
        //    GPU would otherwise idle)
        act_slots.clear();
        zero_iter.clear();
        while (!staged.empty()) {
            Staged &g = staged.front();
            hipError_t q = hipEventQuery(g.copied);
            if (q == hipErrorNotReady && active == 0 && act_slots.empty())
                q = hipEventSynchronize(g.copied);
            if (q == hipErrorNotReady)
                break;
            if (q != hipSuccess) {
                clk::set_error("clk_serve_run: admission copy failed");
                fail(CLK_ERR_HIP);
                break;
            }
            if (staged.size() == 1 || staged[1].copied != g.copied)
                (void)hipEventDestroy(g.copied);  // the round's last job
            BJob &J = b->jobs[g.slot];
            J.state = b->staged_state[g.slot];
            b->h_rec[g.slot] = J;
            act_slots.push_back((uint32_t)g.slot);
            host_iter[g.slot] = 0;
            in_step[g.slot] = 0;
            solo_ran[g.slot] = 0;
            if (J.state == CLK_JOB_ACTIVE)
                ++active;
            else  // zero iterations: done at once
                zero_iter.push_back(g);
            staged.pop_front();
        }
        if (rc != CLK_OK)
            break;
        if (!act_slots.empty()) {  // records and first tables, stream-ordered before the next iteration
            if (act_slots.size() > act_cap) {
                if (h_act)
                    (void)hipHostFree(h_act);
                h_act = nullptr;
                act_cap = std::max(act_slots.size(), 2 * act_cap);
                if (hipHostMalloc((void **)&h_act, act_cap * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
                    hipHostGetDevicePointer((void **)&h_act_dev, h_act, 0) != hipSuccess) {
                    clk::set_error("clk_serve_run: pinned buffer allocation failed");
                    fail(CLK_ERR_NOMEM);
                    break;
                }
            }
            std::memcpy(h_act, act_slots.data(), act_slots.size() * sizeof(uint32_t));
            if (b->fp16)
                hipLaunchKernelGGL(clk::activate_kernel<true>, dim3((unsigned)act_slots.size()), dim3(256), 0, st,
                                   h_act_dev, b->h_rec_dev, b->d_jobs);
            else
                hipLaunchKernelGGL(clk::activate_kernel<false>, dim3((unsigned)act_slots.size()), dim3(256), 0, st,
                                   h_act_dev, b->h_rec_dev, b->d_jobs);
            if (hipGetLastError() != hipSuccess) {
                clk::set_error("clk_serve_run: activation failed");
                fail(CLK_ERR_HIP);
                break;
            }
            b->dirty = true;
            for (const Staged &g : zero_iter) {
                clk_serve_job &j = jobs[g.job];
                const int r2 = clk_batch_take(b, g.slot, j.out_centroids, j.out_labels, j.out_masses,
                                              &j.iterations, &j.shift, st);
                if (r2 != CLK_OK) {
                    fail(r2);
                    break;
                }
                j.finished = seconds_since(t0);
                slot_job[g.slot] = -1;
                ++finished;
            }
            if (rc != CLK_OK)
                break;
        }
        // 4. the next iteration of every resident job
        if (active > 0) {
            max_active = std::max(max_active, active);
            int32_t lone = -1;
            if (active == 1 && staged.empty() && (next >= count || jobs[next].arrival > seconds_since(t0)))
                for (uint32_t s = 0; s < slots && lone < 0; ++s)
                    if (b->jobs[s].state == CLK_JOB_ACTIVE && b->jobs[s].n <= kSoloMaxPoints)
                        lone = (int32_t)s;
            int r = CLK_ERR_UNSUPPORTED;
            if (lone >= 0) {
                const BJob &J = b->jobs[lone];  // host record: buffers and sizes
                const uint32_t it = host_iter[lone];
                const clk::GridRef g{J.gmax, clk::grid_bits(J.n), 1u};  // computed from J.pts at admission
                r = clk::lloyd_step_with_grid(b->ctx, &solo_cfg, J.d, J.n, J.k, J.pts, J.cen[it & 1],
                                              J.cen[(it & 1) ^ 1], J.part, J.labels, J.masses, solo_shift, g, st);
                if (r == CLK_OK) {
                    hipLaunchKernelGGL(clk::solo_finish_kernel, dim3(1), dim3(64), 0, st, b->d_jobs, (uint32_t)lone,
                                       solo_shift, b->h_status_dev);
                    r = hipGetLastError() == hipSuccess ? CLK_OK : CLK_ERR_HIP;
                    solo_ran[lone] = 1;
                    ++solo_steps;
                }
            }
            if (r == CLK_ERR_UNSUPPORTED) {  // the batch step
                r = CLK_OK;
                for (uint32_t s = 0; s < slots && r == CLK_OK; ++s) {
                    if (!solo_ran[s])
                        continue;
                    solo_ran[s] = 0;
                    if (b->jobs[s].state == CLK_JOB_ACTIVE && b->jobs[s].kind == 1) {
                        if (b->fp16)
                            hipLaunchKernelGGL(clk::batch_build_table_kernel<true>, dim3(1), dim3(256), 0, st,
                                               b->d_jobs, s);
                        else
                            hipLaunchKernelGGL(clk::batch_build_table_kernel<false>, dim3(1), dim3(256), 0, st,
                                               b->d_jobs, s);
                        if (hipGetLastError() != hipSuccess)
                            r = CLK_ERR_HIP;
                    }
                }
                if (r == CLK_OK)
                    r = clk_batch_step(b, st);
            }
            for (uint32_t s = 0; s < slots; ++s)
                in_step[s] = b->jobs[s].state == CLK_JOB_ACTIVE;
            if (r != CLK_OK || hipEventRecord(step_done, st) != hipSuccess) {
                fail(r != CLK_OK ? r : CLK_ERR_HIP);
                break;
            }
            inflight = true;
            ++steps;
        } else if (staged.empty() && next < count && done_slots.empty()) {
            // idle until the next arrival
            const double wait = jobs[next].arrival - seconds_since(t0);
            if (wait > 0.0)
                std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
        }
        // 5. results of the jobs that finished in the previous iteration:
        //    copies into pinned staging on their own stream, one
        //    synchronization, then to the requests' buffers; only then are
        //    their slots and arena ranges free again
        if (!done_slots.empty()) {
            size_t need = 0;
            for (uint32_t s : done_slots) {
                const BJob &J = b->jobs[s];
                need += (size_t)J.d * J.k * 4 + (size_t)J.k * 4 + (jobs[slot_job[s]].out_labels ? J.n * 4 : 0);
            }
            const size_t ndesc = 3 * done_slots.size();
            if (need > res_cap || ndesc > desc_cap) {
                if (res)
                    (void)hipHostFree(res);
                if (descs)
                    (void)hipHostFree(descs);
                res = nullptr;
                descs = nullptr;
                res_cap = std::max(need, 2 * res_cap);
                desc_cap = std::max(ndesc, 2 * desc_cap);
                if (hipHostMalloc((void **)&res, res_cap, hipHostMallocDefault) != hipSuccess ||
                    hipHostMalloc((void **)&descs, desc_cap * sizeof(clk::GatherDesc), hipHostMallocDefault) !=
                        hipSuccess ||
                    hipHostGetDevicePointer((void **)&res_dev, res, 0) != hipSuccess ||
                    hipHostGetDevicePointer((void **)&descs_dev, descs, 0) != hipSuccess) {
                    clk::set_error("clk_serve_run: pinned result buffer allocation failed");
                    fail(CLK_ERR_NOMEM);
                    break;
                }
            }
            size_t at = 0, nd = 0;
            for (uint32_t s : done_slots) {
                const BJob &J = b->jobs[s];
                const size_t cb = (size_t)J.d * J.k * 4, mb = (size_t)J.k * 4;
                const size_t lb = jobs[slot_job[s]].out_labels ? J.n * 4 : 0;
                const float *cen = J.cen[done_status[s].iter & 1];
                descs[nd++] = {reinterpret_cast<const uint32_t *>(cen), reinterpret_cast<uint32_t *>(res_dev + at),
                               cb / 4};
                descs[nd++] = {J.masses, reinterpret_cast<uint32_t *>(res_dev + at + cb), mb / 4};
                if (lb)
                    descs[nd++] = {J.labels, reinterpret_cast<uint32_t *>(res_dev + at + cb + mb), lb / 4};
                at += cb + mb + lb;
            }
            hipLaunchKernelGGL(clk::gather_kernel, dim3((unsigned)nd), dim3(256), 0, rs, descs_dev);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(rs) != hipSuccess) {
                clk::set_error("clk_serve_run: result copy failed");
                fail(CLK_ERR_HIP);
                break;
            }
            at = 0;
            const double now2 = seconds_since(t0);
            for (uint32_t s : done_slots) {
                clk_serve_job &j = jobs[slot_job[s]];
                const BJob &J = b->jobs[s];
                const size_t cb = (size_t)J.d * J.k * 4, mb = (size_t)J.k * 4, lb = j.out_labels ? J.n * 4 : 0;
                if (j.out_centroids)
                    std::memcpy(j.out_centroids, res + at, cb);
                if (j.out_masses)
                    std::memcpy(j.out_masses, res + at + cb, mb);
                if (lb)
                    std::memcpy(j.out_labels, res + at + cb + mb, lb);
                at += cb + mb + lb;
                j.iterations = done_status[s].iter;
                j.shift = done_status[s].shift;
                j.finished = now2;
                // the device record stays DONE (no kernel touches it) until the
                // slot is activated again
                b->jobs[s].state = CLK_JOB_FREE;
                slot_job[s] = -1;
                ++finished;
            }
            done_slots.clear();
        }
    }
    if (inflight)
        (void)hipEventSynchronize(step_done);
    for (size_t i = 0; i < staged.size(); ++i) {  // a round's jobs share one event
        (void)hipEventSynchronize(staged[i].copied);
        if (i + 1 == staged.size() || staged[i + 1].copied != staged[i].copied)
            (void)hipEventDestroy(staged[i].copied);
        b->jobs[staged[i].slot].state = CLK_JOB_FREE;
    }
    if (rc != CLK_OK)  // leave the batch empty
        for (uint32_t s = 0; s < slots; ++s)
            if (b->jobs[s].state != CLK_JOB_FREE)
                (void)clk_batch_take(b, (int32_t)s, nullptr, nullptr, nullptr, nullptr, nullptr, st);
    (void)hipStreamSynchronize(cs);
    (void)hipStreamDestroy(cs);
    (void)hipStreamSynchronize(rs);
    (void)hipStreamDestroy(rs);
    if (step_done)
        (void)hipEventDestroy(step_done);
    if (solo_shift)
        (void)hipFree(solo_shift);
    if (res)
        (void)hipHostFree(res);
    if (descs)
        (void)hipHostFree(descs);
    if (h_act)
        (void)hipHostFree(h_act);
    CLK_TRY(rc);
    if (stats) {
        latency_stats(jobs, count, stats);
        stats->steps = steps;
        stats->max_active = max_active;
        stats->solo_steps = solo_steps;
    }
    return CLK_OK;
}

extern "C" int clk_serve_serial(clk_context *ctx, clk_serve_job *jobs, uint32_t count, clk_serve_stats *stats,
                                void *stream)
{
    CLK_TRY(clk::enter(ctx));
    CLK_TRY(check_serve_jobs(jobs, count));
    hipStream_t st = clk::as_stream(stream);
    uint64_t maxn = 0, maxdn = 0, maxdk = 0, maxk = 0;
    for (uint32_t i = 0; i < count; ++i) {
        maxn = std::max<uint64_t>(maxn, jobs[i].n);
        if (!jobs[i].device_points)
            maxdn = std::max<uint64_t>(maxdn, jobs[i].n * jobs[i].d);
        maxdk = std::max<uint64_t>(maxdk, (uint64_t)jobs[i].d * jobs[i].k);
        maxk = std::max<uint64_t>(maxk, jobs[i].k);
    }
    float *pts = nullptr, *cen[2] = {nullptr, nullptr};
    double *sums = nullptr, *shift_d = nullptr;
    uint32_t *labels = nullptr, *masses = nullptr;
    int rc = CLK_OK;
    if (hipMalloc(&pts, std::max<uint64_t>(maxdn, 1) * 4) != hipSuccess || hipMalloc(&cen[0], maxdk * 4) != hipSuccess ||
        hipMalloc(&cen[1], maxdk * 4) != hipSuccess || hipMalloc(&sums, maxdk * 8) != hipSuccess ||
        hipMalloc(&labels, maxn * 4) != hipSuccess || hipMalloc(&masses, maxk * 4) != hipSuccess ||
        hipMalloc(&shift_d, 8) != hipSuccess) {
        clk::set_error("clk_serve_serial: device allocation failed");
        rc = CLK_ERR_NOMEM;
    }
    clk_kernel_config cfg;
    if (rc == CLK_OK)
        rc = clk_default_config(ctx, CLK_STAGE_FUSED, CLK_FUSED_CLUSTER_MERGE, &cfg);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < count && rc == CLK_OK; ++i) {
        clk_serve_job &j = jobs[i];
        while (seconds_since(t0) < j.arrival)
            std::this_thread::sleep_for(std::chrono::duration<double>(
                std::min(0.001, j.arrival - seconds_since(t0))));
        j.admitted = seconds_since(t0);
        const float *jp = j.device_points ? j.points : pts;
        hipError_t e = j.device_points ? hipSuccess
                                       : hipMemcpyAsync(pts, j.points, j.n * j.d * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(cen[0], j.centroids, (size_t)j.d * j.k * 4, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            clk::set_error(std::string("clk_serve_serial: ") + hipGetErrorString(e));
            rc = CLK_ERR_HIP;
            break;
        }
        // the job's sum grid, pinned for its iterations (the batch computes
        // the same at admission)
        rc = clk_sum_grid_from_points(ctx, j.d, j.n, jp, st);
        if (rc != CLK_OK)
            break;
        uint32_t it = 0;
        double shift = 0.0;
        int cur = 0;
        while (it < j.max_iterations) {
            rc = clk_lloyd_step(ctx, &cfg, j.d, j.n, j.k, jp, cen[cur], cen[cur ^ 1], sums, labels, masses,
                                j.tolerance > 0.0 ? shift_d : nullptr, st);
            if (rc != CLK_OK)
                break;
            cur ^= 1;
            ++it;
            if (j.tolerance > 0.0) {  // the convergence test of every iteration
                e = hipMemcpyAsync(&shift, shift_d, 8, hipMemcpyDeviceToHost, st);
                if (e == hipSuccess)
                    e = hipStreamSynchronize(st);
                if (e != hipSuccess) {
                    clk::set_error(std::string("clk_serve_serial: ") + hipGetErrorString(e));
                    rc = CLK_ERR_HIP;
                    break;
                }
                if (shift < j.tolerance)
                    break;
            }
        }
        if (rc != CLK_OK)
            break;
        e = hipStreamSynchronize(st);
        if (e == hipSuccess && j.out_centroids)
            e = hipMemcpy(j.out_centroids, cen[cur], (size_t)j.d * j.k * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && j.out_labels)
            e = hipMemcpy(j.out_labels, labels, j.n * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && j.out_masses)
            e = hipMemcpy(j.out_masses, masses, (size_t)j.k * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            clk::set_error(std::string("clk_serve_serial: ") + hipGetErrorString(e));
            rc = CLK_ERR_HIP;
            break;
        }
        j.iterations = it;
        j.shift = shift;
        j.finished = seconds_since(t0);
    }
    (void)hipStreamSynchronize(st);
    (void)clk_sum_grid_clear(ctx);
    for (void *p : {(void *)pts, (void *)cen[0], (void *)cen[1], (void *)sums, (void *)labels, (void *)masses,
                    (void *)shift_d})
        if (p)
            (void)hipFree(p);
    CLK_TRY(rc);
    if (stats) {
        latency_stats(jobs, count, stats);
        stats->steps = 0;
        stats->max_active = 1;
        stats->solo_steps = 0;
    }
    return CLK_OK;
}

extern "C" int clk_batch_autotune(clk_batch *b, uint32_t *chosen, void *stream)
{
    CLK_REQUIRE(b, CLK_ERR_INVALID_ARGUMENT, "clk_batch_autotune: null batch");
    CLK_TRY(clk::enter(b->ctx));
    hipStream_t st = clk::as_stream(stream);
    CLK_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<BJob> dev(b->jobs.size());
    CLK_HIP_CHECK(hipMemcpy(dev.data(), b->d_jobs, dev.size() * sizeof(BJob), hipMemcpyDeviceToHost));
    // (the sums are exact: re-tiling changes no result, before or after a
    // job's first iteration; the dry passes' labels are those the next
    // iteration computes anyway)
    uint32_t active = 0;
    for (size_t s = 0; s < dev.size(); ++s)
        if (b->jobs[s].state == CLK_JOB_ACTIVE)
            ++active;
    CLK_REQUIRE(active > 0, CLK_ERR_INVALID_ARGUMENT, "clk_batch_autotune: no active jobs to time");
    // re-tile every active job (records on the host and the device)
    auto retile = [&](uint32_t tw) -> int {
        for (size_t s = 0; s < b->jobs.size(); ++s) {
            BJob &J = b->jobs[s];
            if (J.state != CLK_JOB_ACTIVE)
                continue;
            J.tile_pts = tile_points(tw, J.d, J.k);
            J.tile_base = J.tile_pts;
            J.ntiles = (uint32_t)clk::div_up(J.n, J.tile_pts);
            dev[s].tile_pts = J.tile_pts;
            dev[s].tile_base = J.tile_base;
            dev[s].ntiles = J.ntiles;
        }
        CLK_HIP_CHECK(hipMemcpy(b->d_jobs, dev.data(), dev.size() * sizeof(BJob), hipMemcpyHostToDevice));
        b->dirty = true;
        return CLK_OK;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CLK_HIP_CHECK(hipEventCreate(&e0));
    CLK_HIP_CHECK(hipEventCreate(&e1));
    b->dry = true;
    uint32_t best = b->tile_work;
    float best_ms = 0.0f;
    int rc = CLK_OK;
    for (uint32_t tw : kTileWorkCandidates) {
        rc = retile(tw);
        if (rc == CLK_OK)
            rc = rebuild(b, st);
        if (rc == CLK_OK)
            rc = launch_step(b, st);  // warm-up
        if (rc != CLK_OK)
            break;
        float ms = 0.0f;
        if (hipEventRecord(e0, st) != hipSuccess) {
            rc = CLK_ERR_HIP;
            break;
        }
        for (int i = 0; i < 3 && rc == CLK_OK; ++i)
            rc = launch_step(b, st);
        if (rc != CLK_OK || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = rc != CLK_OK ? rc : CLK_ERR_HIP;
            break;
        }
        if (best_ms == 0.0f || ms < best_ms) {
            best_ms = ms;
            best = tw;
        }
    }
    b->dry = false;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == CLK_ERR_HIP)
        clk::set_error("clk_batch_autotune: timing failed");
    const int rc2 = retile(best);  // also on failure: a consistent tiling
    b->tile_work = best;
    CLK_TRY(rc);
    CLK_TRY(rc2);
    CLK_TRY(rebuild(b, st));
    if (chosen)
        *chosen = best;
    return CLK_OK;
}
