// Shared host/device machinery of the gfx950 Lloyd kernels: error reporting,
// the context (scratch buffers, device limits, per-kernel timing records),
// launch helpers and the device-side distance / argmin / reduction helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <string>
#include <vector>

#include "clkmeans.h"

namespace clk {

// ---------------------------------------------------------------- errors --

void set_error(const std::string &msg);

#define CLK_HIP_CHECK(call)                                                      \
    do {                                                                         \
        hipError_t clk_e_ = (call);                                              \
        if (clk_e_ != hipSuccess) {                                              \
            ::clk::set_error(std::string(#call) + ": " + hipGetErrorString(clk_e_)); \
            return CLK_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)

#define CLK_REQUIRE(cond, code, msg)                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            ::clk::set_error(msg);                                               \
            return code;                                                         \
        }                                                                        \
    } while (0)

#define CLK_TRY(expr)                                                            \
    do {                                                                         \
        int clk_rc_ = (expr);                                                    \
        if (clk_rc_ != CLK_OK)                                                   \
            return clk_rc_;                                                      \
    } while (0)

// --------------------------------------------------------------- context --

struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct TimingRecord {
    std::string name;
    int32_t iteration;
    hipEvent_t start;
    hipEvent_t stop;
};

}  // namespace clk

struct clk_context {
    int device = 0;
    int num_cus = 0;
    size_t lds_per_block = 0;     // bytes a work-group may allocate
    int max_threads_per_block = 0;
    clk::Scratch part_sums;       // per-block FP64 partial sums
    clk::Scratch part_masses;     // per-block uint32 partial masses
    clk::Scratch mfma_tables;     // 16-bit centroid operand table, norms, bounds
    clk::Scratch mfma_queue;      // point ids queued for exact refinement
    clk::Scratch mfma_hint;       // largest per-wave score offset K of the last MFMA labeling
    uint32_t *mfma_qlen = nullptr;  // device queue length inside mfma_tables
    clk::Scratch mfma_pin;        // pinned 16-bit operand copy of points (clk_labeling_pin_points)
    const void *mfma_pin_src = nullptr;  // the points it was built from (null: none)
    uint64_t mfma_pin_d = 0, mfma_pin_n = 0;
    bool mfma_pin_fp16 = false;
    clk::Scratch grid;            // sum grid: per-feature max |x| bits [d] (clk_grid.hpp)
    bool grid_pinned = false;     // grid holds a caller-pinned grid for grid_d features
    char *noalloc_scratch = nullptr;  // this device's static scratch (clk_lloyd_step_noalloc)
    uint64_t grid_d = 0, grid_n = 0;
    const void *grid_src = nullptr;  // the points a grid was pinned from (null: from maxima)
    int mfma_subnormals = -1;     // bit 0 FP16, bit 1 BF16: MFMA keeps subnormal inputs (-1 unknown)
    bool profiling = false;
    int32_t iteration = -1;
    std::vector<clk::TimingRecord> records;
    std::vector<hipEvent_t> event_pool;
};

namespace clk {

// Make the context's device current; checks arguments common to all ops.
int enter(clk_context *ctx);

// Grow a scratch buffer to at least `bytes` (hipMalloc; never shrinks).
int ensure_scratch(Scratch &s, size_t bytes);

// Bracket one kernel launch with timing events when profiling is on.
struct KernelTimer {
    KernelTimer(clk_context *ctx, const char *name, hipStream_t stream);
    ~KernelTimer();
    clk_context *ctx;
    hipStream_t stream;
    int64_t rec = -1;
};

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Points per work-item step of the streaming passes: the requested width
// (1, 2 or 4; 8 and above give 4, anything else 4).  Any n and any 4-byte
// aligned buffers take it: the vector accesses need only 4-byte alignment
// (rows start at f*n), and the kernels finish the n % width tail points one
// at a time.
int pick_vector_width(uint64_t requested);

// Persistent-style grid: blocks_per_cu resident blocks on every CU, capped by
// the work available.
uint64_t auto_grid(clk_context *ctx, const void *kernel, int block,
                   size_t dyn_lds, uint64_t work_items);

// ----------------------------------------------------------- device side --

constexpr int kWave = 64;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// The same vectors at 4-byte alignment, for point rows and labels in global
// memory: a row f starts at f*n floats, 16-byte aligned only when 4 divides
// n.  gfx950 global loads and stores of 2 or 4 dwords take any 4-byte
// aligned address (the HSA runtime runs the GPU in unaligned access mode),
// so the accesses stay 64 or 128 bits wide for every n.
typedef float gf32x4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float gf32x2 __attribute__((ext_vector_type(2), aligned(4)));
typedef uint32_t gu32x4 __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t gu32x2 __attribute__((ext_vector_type(2), aligned(4)));

// Streaming (read-once / write-once) accesses: non-temporal loads and stores
// keep the points and labels from displacing reused data in L2 / MALL; on
// gfx950 they measured ~8% faster for the fused pass's read+write mix.
template <typename V>
__device__ __forceinline__ V stream_load(const V *p) { return __builtin_nontemporal_load(p); }
template <typename V>
__device__ __forceinline__ void stream_store(V v, V *p) { __builtin_nontemporal_store(v, p); }

// Point x (VEC consecutive points of DMAX features; features >= d read as 0)
// against every centroid staged in LDS as cen[c * DMAX + f] (padded with 0).
// Distances are the reference's sequential fma over f = 0..d-1 of (x_f-c_f)^2
// from 0 (lloyd_fused_feature_sum.cl:213-239); padded features add
// fma(0, 0, dist) == dist exactly.  Strict '<' against a running minimum
// starting at FLT_MAX; label 0 when nothing is smaller (kmeans_naive.cpp:49-63).
// STRIDE: floats between consecutive centroid rows (DMAX, or more where
// rows are padded against LDS bank conflicts).
template <int DMAX, int VEC, int STRIDE = DMAX>
__device__ __forceinline__ void argmin_lds(const float (&x)[DMAX][VEC],
                                           const float *__restrict__ cen,
                                           uint32_t k, uint32_t (&lab)[VEC])
{
    float best[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        best[v] = FLT_MAX;
        lab[v] = 0;
    }
    for (uint32_t c = 0; c < k; ++c) {
        float cc[DMAX];
#pragma unroll
        for (int f = 0; f < DMAX; ++f)
            cc[f] = cen[c * STRIDE + f];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            float t = x[0][v] - cc[0];
            float dist = t * t;
#pragma unroll
            for (int f = 1; f < DMAX; ++f) {
                t = x[f][v] - cc[f];
                dist = __builtin_fmaf(t, t, dist);
            }
            bool lt = dist < best[v];
            best[v] = lt ? dist : best[v];
            lab[v] = lt ? c : lab[v];
        }
    }
}

// As argmin_lds over centroids c0 .. c0+cn-1 staged in LDS as
// cen[(c - c0) * DMAX + f], continuing a running minimum (best, lab) over the
// centroids before c0: visiting chunks in increasing order with strict '<'
// gives the same label as one pass.
template <int DMAX, int VEC>
__device__ __forceinline__ void argmin_lds_chunk(const float (&x)[DMAX][VEC],
                                                 const float *__restrict__ cen, uint32_t c0,
                                                 uint32_t cn, float (&best)[VEC],
                                                 uint32_t (&lab)[VEC])
{
    for (uint32_t c = 0; c < cn; ++c) {
        float cc[DMAX];
#pragma unroll
        for (int f = 0; f < DMAX; ++f)
            cc[f] = cen[c * DMAX + f];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            float t = x[0][v] - cc[0];
            float dist = t * t;
#pragma unroll
            for (int f = 1; f < DMAX; ++f) {
                t = x[f][v] - cc[f];
                dist = __builtin_fmaf(t, t, dist);
            }
            bool lt = dist < best[v];
            best[v] = lt ? dist : best[v];
            lab[v] = lt ? c0 + c : lab[v];
        }
    }
}

// Load VEC consecutive points p .. p+VEC-1 (all < n) of DMAX features;
// features >= d are 0.  EXACT: d == DMAX (no per-feature guards).
template <int DMAX, int VEC, bool EXACT = false>
__device__ __forceinline__ void load_points(const float *__restrict__ pts,
                                            uint64_t n, uint32_t d, uint64_t p,
                                            float (&x)[DMAX][VEC])
{
    // Row bases are work-group uniform (SGPRs); when every row offset fits in
    // 32 bits the per-lane address is one 32-bit offset (saddr addressing).
    const bool off32 = n < (1ull << 30);
    const uint32_t p32 = (uint32_t)p;
#pragma unroll
    for (int f = 0; f < DMAX; ++f) {
        if (EXACT || f < (int)d) {
            const float *rowbase = pts + (uint64_t)f * n;
            const float *row = off32 ? rowbase + p32 : rowbase + p;
            if constexpr (VEC == 1) {
                x[f][0] = stream_load(row);
            } else if constexpr (VEC == 2) {
                gf32x2 a = stream_load(reinterpret_cast<const gf32x2 *>(row));
                x[f][0] = a.x; x[f][1] = a.y;
            } else if constexpr (VEC == 4) {
                gf32x4 a = stream_load(reinterpret_cast<const gf32x4 *>(row));
                x[f][0] = a.x; x[f][1] = a.y; x[f][2] = a.z; x[f][3] = a.w;
            } else {
                gf32x4 a = stream_load(reinterpret_cast<const gf32x4 *>(row));
                gf32x4 b = stream_load(reinterpret_cast<const gf32x4 *>(row + 4));
                x[f][0] = a.x; x[f][1] = a.y; x[f][2] = a.z; x[f][3] = a.w;
                x[f][4] = b.x; x[f][5] = b.y; x[f][6] = b.z; x[f][7] = b.w;
            }
        } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                x[f][v] = 0.0f;
        }
    }
}

template <int VEC>
__device__ __forceinline__ void store_labels(uint32_t *__restrict__ labels,
                                             uint64_t p, const uint32_t (&lab)[VEC])
{
    uint32_t *dst = labels + p;
    if constexpr (VEC == 1) {
        stream_store(lab[0], dst);
    } else if constexpr (VEC == 2) {
        stream_store(gu32x2{lab[0], lab[1]}, reinterpret_cast<gu32x2 *>(dst));
    } else if constexpr (VEC == 4) {
        stream_store(gu32x4{lab[0], lab[1], lab[2], lab[3]}, reinterpret_cast<gu32x4 *>(dst));
    } else {
        stream_store(gu32x4{lab[0], lab[1], lab[2], lab[3]}, reinterpret_cast<gu32x4 *>(dst));
        stream_store(gu32x4{lab[4], lab[5], lab[6], lab[7]}, reinterpret_cast<gu32x4 *>(dst + 4));
    }
}

template <int VEC>
__device__ __forceinline__ void load_labels(const uint32_t *__restrict__ labels,
                                            uint64_t p, uint32_t (&lab)[VEC])
{
    const uint32_t *src = labels + p;
    if constexpr (VEC == 1) {
        lab[0] = src[0];
    } else if constexpr (VEC == 2) {
        gu32x2 a = *reinterpret_cast<const gu32x2 *>(src);
        lab[0] = a.x; lab[1] = a.y;
    } else if constexpr (VEC == 4) {
        gu32x4 a = *reinterpret_cast<const gu32x4 *>(src);
        lab[0] = a.x; lab[1] = a.y; lab[2] = a.z; lab[3] = a.w;
    } else {
        gu32x4 a = *reinterpret_cast<const gu32x4 *>(src);
        gu32x4 b = *reinterpret_cast<const gu32x4 *>(src + 4);
        lab[0] = a.x; lab[1] = a.y; lab[2] = a.z; lab[3] = a.w;
        lab[4] = b.x; lab[5] = b.y; lab[6] = b.z; lab[7] = b.w;
    }
}

// load_labels with non-temporal loads (labels read once per pass).
template <int VEC>
__device__ __forceinline__ void stream_load_labels(const uint32_t *__restrict__ labels,
                                                   uint64_t p, uint32_t (&lab)[VEC])
{
    static_assert(VEC == 1 || VEC == 2 || VEC == 4, "1, 2 or 4 labels");
    const uint32_t *src = labels + p;
    if constexpr (VEC == 1) {
        lab[0] = stream_load(src);
    } else if constexpr (VEC == 2) {
        gu32x2 a = stream_load(reinterpret_cast<const gu32x2 *>(src));
        lab[0] = a.x; lab[1] = a.y;
    } else {
        gu32x4 a = stream_load(reinterpret_cast<const gu32x4 *>(src));
        lab[0] = a.x; lab[1] = a.y; lab[2] = a.z; lab[3] = a.w;
    }
}

// Keep-label threshold of centroid a for the bounded fused pass.
//
// s2 = min over c != a of ||c_a - c||^2 (FP64 from the FP32 centroids;
// NaN distances skipped).  With m = 2^-10 and thr = s2 (1-m) / (2+m)^2
// (rounded down to float, capped at 1e30), a point x whose FP32 distance
// D(x, c_a) -- the reference's fma sequence -- is <= thr gets label a from
// the full scan too:
//   the FP32 sum of d squares of rounded differences is within a relative
//   (d+3)*2^-24 of the exact r^2 = ||x - c_a||^2 (plus at most d*2^-126
//   absolute from underflow), so r < sqrt(s2) / (2+m); the triangle
//   inequality gives ||x - c|| >= sqrt(s2) - r > sqrt(s2) (1+m) / (2+m) for
//   every other centroid, hence D(x, c) > thr (1-(d+3)2^-24)(1+m)^2/(1-m)
//   - d*2^-126 > thr >= D(x, c_a): c_a is the unique strict minimum and
//   below FLT_MAX, which is what the strict-'<' scan from FLT_MAX returns
//   (kmeans_naive.cpp:49-63).  The margin m dwarfs the rounding terms for
//   d <= 2^12 (the FP64 arithmetic here adds a few 2^-53); thresholds below
//   1e-20 (duplicate or nearly coincident centroids) return -1: never keep.
__device__ __forceinline__ float keep_threshold(double s2)
{
    constexpr double m = 1.0 / 1024.0;
    if (!(s2 > 0.0))
        return -1.0f;
    double t = s2 * (1.0 - m) / ((2.0 + m) * (2.0 + m));
    if (t > 1e30)
        t = 1e30;  // also s2 = inf: k == 1, or every other centroid is non-finite
    if (t < 1e-20)
        return -1.0f;
    return __double2float_rd(t);
}

// Stage centroids [d][k] (global) into LDS as [c][DMAX], zero-padded.
template <int DMAX, int STRIDE = DMAX>
__device__ __forceinline__ void stage_centroids(const float *__restrict__ cen,
                                                uint32_t d, uint32_t k,
                                                float *__restrict__ lds)
{
    for (uint32_t i = threadIdx.x; i < k * DMAX; i += blockDim.x) {
        uint32_t c = i / DMAX, f = i % DMAX;
        lds[c * STRIDE + f] = (f < d) ? cen[(uint64_t)f * k + c] : 0.0f;
    }
}

// keep_threshold of every centroid staged in LDS as [c][DMAX] (padding
// features are 0 in every row, so they add nothing).  Call after the
// centroids are visible to the work-group; thr[] needs a barrier after.
template <int DMAX>
__device__ __forceinline__ void stage_keep_thresholds(const float *__restrict__ cen_lds,
                                                      uint32_t k, float *__restrict__ thr)
{
    for (uint32_t a = threadIdx.x; a < k; a += blockDim.x) {
        double s2 = __builtin_inf();
        for (uint32_t c = 0; c < k; ++c) {
            if (c == a)
                continue;
            double acc = 0.0;
#pragma unroll
            for (int f = 0; f < DMAX; ++f) {
                const double t = (double)cen_lds[a * DMAX + f] - (double)cen_lds[c * DMAX + f];
                acc = __builtin_fma(t, t, acc);
            }
            s2 = acc < s2 ? acc : s2;  // a NaN distance never lowers it
        }
        thr[a] = keep_threshold(s2);
    }
}

// Add into a work-item-private LDS cell: plain read-add-write (no other
// work-item touches the address).  On gfx950 this is ~5x faster than
// ds_add_f32 for the fused pass.
__device__ __forceinline__ void lds_acc(float *p, float v) { *p = *p + v; }
__device__ __forceinline__ void lds_acc(uint32_t *p, uint32_t v) { *p = *p + v; }

// LDS float / uint atomics without return value (ds_add_f32 / ds_add_u32),
// for cells shared between work-items.
__device__ __forceinline__ void lds_add(float *p, float v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// End-of-block merge of work-item-private LDS tiles laid out [value][cols]:
// one thread sums one value's `cols` lane entries in FP64 in a fixed order.
// The start column is rotated by `rot` (the thread's value index) so the
// threads of a wave read different LDS banks at each step.
__device__ __forceinline__ double lds_row_sum_f32(const float *__restrict__ row,
                                                  uint32_t cols, uint32_t rot)
{
    double s0 = 0.0, s1 = 0.0;
    uint32_t j = rot % cols;
    for (uint32_t i = 0; i < cols; ++i) {
        double v = (double)row[j];
        if (i & 1) s1 += v; else s0 += v;
        j = (j + 1 == cols) ? 0 : j + 1;
    }
    return s0 + s1;
}

__device__ __forceinline__ uint32_t lds_row_sum_u32(const uint32_t *__restrict__ row,
                                                    uint32_t cols, uint32_t rot)
{
    uint32_t s = 0;
    uint32_t j = rot % cols;
    for (uint32_t i = 0; i < cols; ++i) {
        s += row[j];
        j = (j + 1 == cols) ? 0 : j + 1;
    }
    return s;
}

}  // namespace clk
