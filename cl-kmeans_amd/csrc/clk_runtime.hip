// Context, error reporting, per-kernel timing, and the small element-wise /
// reduction kernels of the Lloyd iteration: fill, first_x, matrix add,
// column-vector reduction (ReduceVectorParcol) and the FP64 divide
// (MatrixBinaryOp::row).
#include "clk_common.hpp"
#include "clk_finalize.hpp"
#include "clk_kernels.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace clk {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int enter(clk_context *ctx)
{
    CLK_REQUIRE(ctx != nullptr, CLK_ERR_INVALID_ARGUMENT, "null clk_context");
    CLK_HIP_CHECK(hipSetDevice(ctx->device));
    return CLK_OK;
}

int ensure_scratch(Scratch &s, size_t bytes)
{
    if (s.bytes >= bytes)
        return CLK_OK;
    if (s.ptr) {
        CLK_HIP_CHECK(hipFree(s.ptr));
        s.ptr = nullptr;
        s.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 4096);
    hipError_t e = hipMalloc(&s.ptr, want);
    if (e != hipSuccess) {
        s.ptr = nullptr;
        set_error(std::string("scratch hipMalloc failed: ") + hipGetErrorString(e));
        return CLK_ERR_NOMEM;
    }
    s.bytes = want;
    return CLK_OK;
}

static hipEvent_t take_event(clk_context *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

KernelTimer::KernelTimer(clk_context *c, const char *name, hipStream_t s) : ctx(c), stream(s)
{
    if (!ctx->profiling)
        return;
    hipEvent_t a = take_event(ctx), b = take_event(ctx);
    if (!a || !b)
        return;
    ctx->records.push_back(TimingRecord{name, ctx->iteration, a, b});
    rec = (int64_t)ctx->records.size() - 1;
    (void)hipEventRecord(a, stream);
}

KernelTimer::~KernelTimer()
{
    if (rec >= 0)
        (void)hipEventRecord(ctx->records[(size_t)rec].stop, stream);
}

int pick_vector_width(uint64_t requested)
{
    if (requested == 1 || requested == 2 || requested == 4)
        return (int)requested;
    return 4;
}

uint64_t auto_grid(clk_context *ctx, const void *kernel, int block, size_t dyn_lds,
                   uint64_t work_items)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, dyn_lds) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    uint64_t grid = (uint64_t)per_cu * (uint64_t)ctx->num_cus;
    uint64_t need = div_up(std::max<uint64_t>(work_items, 1), (uint64_t)block);
    return std::max<uint64_t>(1, std::min(grid, need));
}

// ------------------------------------------------------------ kernels ---

namespace {

__global__ void fill_zero_kernel(uint4 *__restrict__ p16, uint64_t n16,
                                 uint8_t *__restrict__ tail, uint64_t ntail)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = i; j < n16; j += stride)
        p16[j] = make_uint4(0, 0, 0, 0);
    for (uint64_t j = i; j < ntail; j += stride)
        tail[j] = 0;
}

__global__ void first_x_kernel(const float *__restrict__ pts, uint64_t n, uint32_t d,
                               uint32_t k, float *__restrict__ cen)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)d * k)
        return;
    uint64_t f = i / k, c = i % k;
    cen[i] = pts[f * n + (c % n)];
}

template <typename T>
__global__ void matrix_add_kernel(T *__restrict__ fst, const T *__restrict__ snd, uint64_t count)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride)
        fst[i] = fst[i] + snd[i];
}

// Column reduction with a fixed order.  A block covers 64 rows with
// G = blockDim.x/64 column groups; thread (lane, g) sums columns
// g, g+G, g+2G, ... of row r0+lane into four interleaved accumulators
// (independent load chains), combines them as (a0+a1)+(a2+a3), and the G
// group sums are added in group order.  `in` and `out` may alias (in-place):
// a block reads and writes only its own rows.
template <typename T>
__device__ __forceinline__ T column_sum(const T *in, uint64_t cols, uint64_t rows, uint64_t r,
                                        uint32_t g, uint32_t G)
{
    T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    uint64_t c = g;
    for (; c + 3 * (uint64_t)G < cols; c += 4 * (uint64_t)G) {
        a0 += in[c * rows + r];
        a1 += in[(c + G) * rows + r];
        a2 += in[(c + 2 * (uint64_t)G) * rows + r];
        a3 += in[(c + 3 * (uint64_t)G) * rows + r];
    }
    for (; c < cols; c += G)
        a0 += in[c * rows + r];
    return (T)((T)(a0 + a1) + (T)(a2 + a3));
}

template <typename T>
__global__ void __launch_bounds__(1024) reduce_cols_kernel(const T *in, uint64_t cols,
                                                           uint64_t rows, T *out,
                                                           int accumulate)
{
    __shared__ T part[1024];
    const uint32_t lane = threadIdx.x % 64, g = threadIdx.x / 64;
    const uint32_t G = blockDim.x / 64;
    const uint64_t r = (uint64_t)blockIdx.x * 64 + lane;
    part[threadIdx.x] = (r < rows) ? column_sum(in, cols, rows, r, g, G) : (T)0;
    __syncthreads();
    if (g == 0 && r < rows) {
        T t = part[lane];
        for (uint32_t j = 1; j < G; ++j)
            t += part[j * 64 + lane];
        out[r] = accumulate ? (T)(out[r] + t) : t;
    }
}

// The fused pass's two partial sets in one launch: rows [0, rows_s) are the
// FP64 sums, rows [rows_s, rows_s + rows_m) the uint32 masses; both are
// accumulated into their outputs.
__global__ void __launch_bounds__(1024) reduce_partials_kernel(
    const double *ps, uint64_t rows_s, double *sums, const uint32_t *pm, uint64_t rows_m,
    uint32_t *masses, uint64_t cols)
{
    __shared__ double part_d[1024];
    __shared__ uint32_t part_u[1024];
    const uint32_t lane = threadIdx.x % 64, g = threadIdx.x / 64;
    const uint32_t G = blockDim.x / 64;
    const uint64_t r = (uint64_t)blockIdx.x * 64 + lane;
    const bool is_sum = r < rows_s;
    const bool is_mass = !is_sum && r < rows_s + rows_m;
    part_d[threadIdx.x] = is_sum ? column_sum(ps, cols, rows_s, r, g, G) : 0.0;
    part_u[threadIdx.x] = is_mass ? column_sum(pm, cols, rows_m, r - rows_s, g, G) : 0u;
    __syncthreads();
    if (g != 0)
        return;
    if (is_sum) {
        double t = part_d[lane];
        for (uint32_t j = 1; j < G; ++j)
            t += part_d[j * 64 + lane];
        sums[r] += t;
    } else if (is_mass) {
        uint32_t t = part_u[lane];
        for (uint32_t j = 1; j < G; ++j)
            t += part_u[j * 64 + lane];
        masses[r - rows_s] += t;
    }
}

// centroids[f,c] = (float)(sums[f,c] / (double)masses[c]); optional max
// centroid shift (FP64) over non-empty clusters.  One block.
__global__ void __launch_bounds__(1024) divide_kernel(const double *__restrict__ sums,
                                                      const uint32_t *__restrict__ masses,
                                                      uint32_t d, uint32_t k,
                                                      float *__restrict__ cen,
                                                      const float *__restrict__ old_cen,
                                                      double *__restrict__ max_shift)
{
    __shared__ double red[1024];
    double local_max = 0.0;
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) {
        const double m = (double)masses[c];
        double sq = 0.0;
        for (uint32_t f = 0; f < d; ++f) {
            const uint64_t i = (uint64_t)f * k + c;
            const float v = (float)(sums[i] / m);
            cen[i] = v;
            if (old_cen) {
                const double diff = (double)v - (double)old_cen[i];
                sq += diff * diff;
            }
        }
        if (masses[c] != 0u)
            local_max = fmax(local_max, sqrt(sq));
    }
    if (!max_shift || !old_cen)
        return;
    red[threadIdx.x] = local_max;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s /= 2) {
        if (threadIdx.x < s)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *max_shift = red[0];
}

}  // namespace

// 64 rows x up to 16 column groups; fewer groups for few columns.
static unsigned reduce_block(uint64_t cols)
{
    unsigned groups = 1;
    while (groups < 16 && (uint64_t)groups * 8 < cols)
        groups *= 2;
    return 64 * groups;
}

int launch_reduce_partials(clk_context *ctx, const double *ps, uint64_t rows_s, double *sums,
                           const uint32_t *pm, uint64_t rows_m, uint32_t *masses, uint64_t cols,
                           hipStream_t stream)
{
    if (rows_s + rows_m == 0)
        return CLK_OK;
    KernelTimer t(ctx, "ReduceVectorParcol", stream);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)div_up(rows_s + rows_m, 64)),
                       dim3(reduce_block(cols)), 0, stream, ps, rows_s, sums, pm, rows_m, masses,
                       cols);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

namespace {

// One work-group of kFinalizeThreads threads.  Thread t sums values
// v = t mod 64 + 64*i over the row slice {g, g+G, ...}, g = t / 64,
// G = 16, with four interleaved accumulators; the 16 slice sums of a value
// are then added in slice order.  Dynamic LDS: d*k doubles + k uint32.
__global__ void __launch_bounds__(kFinalizeThreads) finalize_kernel(FinalizeArgs a)
{
    extern __shared__ __align__(16) double fin_lds[];
    __shared__ double part_d[kFinalizeThreads];
    __shared__ uint32_t part_u[kFinalizeThreads];
    __shared__ unsigned long long s_max;  // bits of a non-negative double
    const uint32_t tid = threadIdx.x, lane = tid % 64, g = tid / 64;
    constexpr uint32_t G = kFinalizeThreads / 64;
    const uint64_t dk = (uint64_t)a.d * a.k;
    const uint64_t V = dk + a.k;
    double *tot_s = fin_lds;
    uint32_t *tot_m = reinterpret_cast<uint32_t *>(tot_s + dk);

    for (uint64_t v0 = 0; v0 < V; v0 += 64) {
        const uint64_t v = v0 + lane;
        double sd = 0.0;
        uint32_t su = 0;
        if (v < dk) {
            // eight independent load chains per thread
            double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            uint32_t r = g;
            for (; r + 7 * G < a.rows; r += 8 * G) {
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc[u] += a.part_sums[(uint64_t)(r + u * G) * dk + v];
            }
            for (; r < a.rows; r += G)
                acc[0] += a.part_sums[(uint64_t)r * dk + v];
            sd = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        } else if (v < V) {
            const uint64_t c = v - dk;
            uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            uint32_t r = g;
            for (; r + 7 * G < a.rows; r += 8 * G) {
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc[u] += a.part_masses[(uint64_t)(r + u * G) * a.k + c];
            }
            for (; r < a.rows; r += G)
                acc[0] += a.part_masses[(uint64_t)r * a.k + c];
            for (int u = 0; u < 8; ++u)
                su += acc[u];
        }
        part_d[tid] = sd;
        part_u[tid] = su;
        __syncthreads();
        if (g == 0 && v < V) {
            if (v < dk) {
                double t = part_d[lane];
                for (uint32_t j = 1; j < G; ++j)
                    t += part_d[j * 64 + lane];
                if (a.mode == 1)
                    a.sums[v] += t;
                else
                    a.sums[v] = t;
                tot_s[v] = t;
            } else {
                uint32_t t = part_u[lane];
                for (uint32_t j = 1; j < G; ++j)
                    t += part_u[j * 64 + lane];
                const uint64_t c = v - dk;
                if (a.mode == 1)
                    a.masses[c] += t;
                else
                    a.masses[c] = t;
                tot_m[c] = t;
            }
        }
        __syncthreads();
    }
    if (a.mode != 2)
        return;
    if (tid == 0)
        s_max = 0ull;
    __syncthreads();
    // FP64 divide, one rounding to float; empty cluster -> 0/0 = NaN.
    double local = 0.0;
    for (uint32_t c = tid; c < a.k; c += kFinalizeThreads) {
        const double m = (double)tot_m[c];
        double sq = 0.0;
        for (uint32_t f = 0; f < a.d; ++f) {
            const uint64_t i = (uint64_t)f * a.k + c;
            const float v = (float)(tot_s[i] / m);
            a.centroids[i] = v;
            if (a.old_centroids) {
                const double diff = (double)v - (double)a.old_centroids[i];
                sq += diff * diff;
            }
        }
        if (tot_m[c] != 0u)
            local = fmax(local, sqrt(sq));
    }
    if (!a.max_shift || !a.old_centroids)
        return;
    // Non-negative doubles order like their bit patterns: integer max.
    atomicMax(&s_max, (unsigned long long)__double_as_longlong(local));
    __syncthreads();
    if (tid == 0)
        *a.max_shift = __longlong_as_double((long long)s_max);
}

}  // namespace

int launch_finalize(clk_context *ctx, const FinalizeArgs &a, hipStream_t stream)
{
    const size_t lds = (size_t)a.d * a.k * sizeof(double) + (size_t)a.k * sizeof(uint32_t);
    CLK_REQUIRE(lds + 12 * 1024 + 16 <= ctx->lds_per_block, CLK_ERR_UNSUPPORTED,
                "finalize: d*k too large for the one-work-group merge");
    KernelTimer t(ctx, a.mode == 2 ? "FinalizeDivide" : "ReduceVectorParcol", stream);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kFinalizeThreads), lds, stream, a);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

int launch_reduce_f64(clk_context *ctx, const double *in, uint64_t cols, uint64_t rows,
                      double *out, bool accumulate, hipStream_t stream)
{
    if (rows == 0)
        return CLK_OK;
    KernelTimer t(ctx, "ReduceVectorParcol", stream);
    hipLaunchKernelGGL(reduce_cols_kernel<double>, dim3((unsigned)div_up(rows, 64)),
                       dim3(reduce_block(cols)), 0, stream, in, cols, rows, out, accumulate ? 1 : 0);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

int launch_reduce_u32(clk_context *ctx, const uint32_t *in, uint64_t cols, uint64_t rows,
                      uint32_t *out, bool accumulate, hipStream_t stream)
{
    if (rows == 0)
        return CLK_OK;
    KernelTimer t(ctx, "ReduceVectorParcol", stream);
    hipLaunchKernelGGL(reduce_cols_kernel<uint32_t>, dim3((unsigned)div_up(rows, 64)),
                       dim3(reduce_block(cols)), 0, stream, in, cols, rows, out, accumulate ? 1 : 0);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

}  // namespace clk

using namespace clk;

// -------------------------------------------------------------- C ABI ---

extern "C" const char *clk_last_error(void) { return g_last_error.c_str(); }

// For the host layer's C entry points (C++ exceptions -> clk_last_error).
extern "C" void clk_set_error_message(const char *msg) { clk::set_error(msg ? msg : ""); }

extern "C" int clk_context_create(int device, clk_context **out)
{
    CLK_REQUIRE(out != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_context_create: null out");
    *out = nullptr;
    int count = 0;
    CLK_HIP_CHECK(hipGetDeviceCount(&count));
    CLK_REQUIRE(device >= 0 && device < count, CLK_ERR_INVALID_ARGUMENT,
                "clk_context_create: no such HIP device");
    CLK_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    CLK_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    clk_context *ctx = new clk_context;
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount;
    ctx->lds_per_block = prop.sharedMemPerBlock;
    ctx->max_threads_per_block = prop.maxThreadsPerBlock;
    const int rc = noalloc_setup(ctx);
    if (rc != CLK_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return CLK_OK;
}

extern "C" int clk_context_destroy(clk_context *ctx)
{
    if (!ctx)
        return CLK_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->part_sums.ptr)
        (void)hipFree(ctx->part_sums.ptr);
    for (clk::Scratch *sc : {&ctx->part_masses, &ctx->mfma_tables, &ctx->mfma_queue, &ctx->mfma_hint, &ctx->mfma_pin, &ctx->grid})
        if (sc->ptr)
            (void)hipFree(sc->ptr);
    for (auto &r : ctx->records) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : ctx->event_pool)
        (void)hipEventDestroy(e);
    delete ctx;
    return CLK_OK;
}

extern "C" int clk_strategy_id(int stage, const char *name)
{
    CLK_REQUIRE(name != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_strategy_id: null name");
    std::string s(name);
    switch (stage) {
    case CLK_STAGE_FUSED:
        if (s == "cluster_merge") return CLK_FUSED_CLUSTER_MERGE;
        if (s == "feature_sum") return CLK_FUSED_FEATURE_SUM;
        if (s == "cluster_merge_bounded") return CLK_FUSED_CLUSTER_MERGE_BOUNDED;
        break;
    case CLK_STAGE_LABELING:
        if (s == "unroll_vector") return CLK_LABELING_UNROLL_VECTOR;
        if (s == "mfma_bf16") return CLK_LABELING_MFMA_BF16;
        if (s == "mfma_fp16") return CLK_LABELING_MFMA_FP16;
        if (s == "unroll_vector_bounded") return CLK_LABELING_UNROLL_VECTOR_BOUNDED;
        break;
    case CLK_STAGE_MASS_UPDATE:
        if (s == "global_atomic") return CLK_MASS_GLOBAL_ATOMIC;
        if (s == "part_global") return CLK_MASS_PART_GLOBAL;
        if (s == "part_local") return CLK_MASS_PART_LOCAL;
        if (s == "part_private") return CLK_MASS_PART_PRIVATE;
        break;
    case CLK_STAGE_CENTROID_UPDATE:
        if (s == "feature_sum") return CLK_CENTROID_FEATURE_SUM;
        if (s == "feature_sum_pardim") return CLK_CENTROID_FEATURE_SUM_PARDIM;
        if (s == "cluster_merge") return CLK_CENTROID_CLUSTER_MERGE;
        break;
    default:
        set_error("clk_strategy_id: unknown stage");
        return CLK_ERR_INVALID_ARGUMENT;
    }
    set_error(s);  // the reference throws std::invalid_argument(strategy)
    return CLK_ERR_INVALID_ARGUMENT;
}

extern "C" int clk_default_config(clk_context *ctx, int stage, int strategy,
                                  clk_kernel_config *cfg)
{
    (void)ctx;
    CLK_REQUIRE(cfg != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_default_config: null cfg");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->stage = stage;
    cfg->strategy = strategy;
    cfg->local_size[0] = 256;
    cfg->local_size[1] = cfg->local_size[2] = 1;
    cfg->global_size[1] = cfg->global_size[2] = 1;
    cfg->global_size[0] = 0;  // persistent grid sized from occupancy
    cfg->vector_length = 4;
    cfg->local_features = 1;
    cfg->thread_features = 1;
    return CLK_OK;
}

extern "C" int clk_profiling_enable(clk_context *ctx, int enable)
{
    CLK_REQUIRE(ctx, CLK_ERR_INVALID_ARGUMENT, "null clk_context");
    ctx->profiling = enable != 0;
    return CLK_OK;
}

extern "C" int clk_profiling_set_iteration(clk_context *ctx, int iteration)
{
    CLK_REQUIRE(ctx, CLK_ERR_INVALID_ARGUMENT, "null clk_context");
    ctx->iteration = iteration;
    return CLK_OK;
}

extern "C" int clk_profiling_clear(clk_context *ctx)
{
    CLK_TRY(enter(ctx));
    for (auto &r : ctx->records) {
        (void)hipEventSynchronize(r.stop);
        ctx->event_pool.push_back(r.start);
        ctx->event_pool.push_back(r.stop);
    }
    ctx->records.clear();
    return CLK_OK;
}

extern "C" int64_t clk_profiling_count(clk_context *ctx)
{
    if (!ctx)
        return CLK_ERR_INVALID_ARGUMENT;
    return (int64_t)ctx->records.size();
}

extern "C" int clk_profiling_get(clk_context *ctx, int64_t i, char *name, size_t name_len,
                                 int32_t *iteration, float *ms)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(i >= 0 && (size_t)i < ctx->records.size(), CLK_ERR_INVALID_ARGUMENT,
                "clk_profiling_get: index out of range");
    TimingRecord &r = ctx->records[(size_t)i];
    if (name && name_len) {
        std::strncpy(name, r.name.c_str(), name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (iteration)
        *iteration = r.iteration;
    if (ms) {
        CLK_HIP_CHECK(hipEventSynchronize(r.stop));
        CLK_HIP_CHECK(hipEventElapsedTime(ms, r.start, r.stop));
    }
    return CLK_OK;
}

extern "C" int clk_fill_zero(clk_context *ctx, void *ptr, uint64_t bytes, void *stream)
{
    CLK_TRY(enter(ctx));
    if (bytes == 0)
        return CLK_OK;
    CLK_REQUIRE(ptr != nullptr, CLK_ERR_INVALID_ARGUMENT, "clk_fill_zero: null pointer");
    hipStream_t s = as_stream(stream);
    uint8_t *p = static_cast<uint8_t *>(ptr);
    uint64_t head = 0;
    if (reinterpret_cast<uintptr_t>(p) % 16 != 0)
        head = bytes;  // unaligned: byte path for everything
    uint64_t n16 = head ? 0 : bytes / 16;
    uint8_t *tail = p + n16 * 16;
    uint64_t ntail = bytes - n16 * 16;
    uint64_t work = std::max(n16, ntail);
    unsigned grid = (unsigned)std::min<uint64_t>(div_up(work, 256), 4096);
    KernelTimer t(ctx, "Fill", s);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(std::max(grid, 1u)), dim3(256), 0, s,
                       reinterpret_cast<uint4 *>(p), n16, tail, ntail);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

extern "C" int clk_init_first_x(clk_context *ctx, uint64_t d, uint64_t n, uint64_t k,
                                const float *points, float *centroids, void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(d >= 1 && n >= 1 && k >= 1, CLK_ERR_SIZE, "first_x: d, n, k must be >= 1");
    CLK_REQUIRE(points && centroids, CLK_ERR_INVALID_ARGUMENT, "first_x: null buffer");
    hipStream_t s = as_stream(stream);
    KernelTimer t(ctx, "FirstX", s);
    hipLaunchKernelGGL(first_x_kernel, dim3((unsigned)div_up(d * k, 256)), dim3(256), 0, s,
                       points, n, (uint32_t)d, (uint32_t)k, centroids);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

template <typename T>
static int matrix_add(clk_context *ctx, uint64_t count, T *fst, const T *snd, void *stream)
{
    CLK_TRY(enter(ctx));
    if (count == 0)
        return CLK_OK;
    CLK_REQUIRE(fst && snd, CLK_ERR_INVALID_ARGUMENT, "matrix_add: null buffer");
    hipStream_t s = as_stream(stream);
    unsigned grid = (unsigned)std::min<uint64_t>(div_up(count, 256), 8192);
    KernelTimer t(ctx, "MatrixBinaryOpMatrix", s);
    hipLaunchKernelGGL(matrix_add_kernel<T>, dim3(grid), dim3(256), 0, s, fst, snd, count);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}

extern "C" int clk_matrix_add_u32(clk_context *ctx, uint64_t count, uint32_t *fst,
                                  const uint32_t *snd, void *stream)
{
    return matrix_add<uint32_t>(ctx, count, fst, snd, stream);
}

extern "C" int clk_matrix_add_f64(clk_context *ctx, uint64_t count, double *fst,
                                  const double *snd, void *stream)
{
    return matrix_add<double>(ctx, count, fst, snd, stream);
}

extern "C" int clk_reduce_vector_parcol_u32(clk_context *ctx, uint64_t num_cols,
                                            uint64_t result_rows, uint32_t *data,
                                            void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(num_cols >= 1, CLK_ERR_SIZE, "reduce_vector_parcol: num_cols must be >= 1");
    CLK_REQUIRE(data || result_rows == 0, CLK_ERR_INVALID_ARGUMENT, "reduce: null data");
    return launch_reduce_u32(ctx, data, num_cols, result_rows, data, false, as_stream(stream));
}

extern "C" int clk_reduce_vector_parcol_f64(clk_context *ctx, uint64_t num_cols,
                                            uint64_t result_rows, double *data,
                                            void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(num_cols >= 1, CLK_ERR_SIZE, "reduce_vector_parcol: num_cols must be >= 1");
    CLK_REQUIRE(data || result_rows == 0, CLK_ERR_INVALID_ARGUMENT, "reduce: null data");
    return launch_reduce_f64(ctx, data, num_cols, result_rows, data, false, as_stream(stream));
}

extern "C" int clk_matrix_divide_row(clk_context *ctx, uint64_t d, uint64_t k,
                                     const double *sums, const uint32_t *masses,
                                     float *centroids, const float *old_centroids,
                                     double *max_shift, void *stream)
{
    CLK_TRY(enter(ctx));
    CLK_REQUIRE(d >= 1 && k >= 1, CLK_ERR_SIZE, "divide: d and k must be >= 1");
    CLK_REQUIRE(d * k < (1ull << 32), CLK_ERR_SIZE, "divide: d*k too large");
    CLK_REQUIRE(sums && masses && centroids, CLK_ERR_INVALID_ARGUMENT, "divide: null buffer");
    CLK_REQUIRE(!old_centroids || old_centroids != centroids, CLK_ERR_INVALID_ARGUMENT,
                "divide: old_centroids must not alias centroids");
    hipStream_t s = as_stream(stream);
    KernelTimer t(ctx, "MatrixBinaryOpRow", s);
    unsigned block = 64;
    while (block < 1024 && block < k)
        block *= 2;  // power of two for the tree reduction
    hipLaunchKernelGGL(divide_kernel, dim3(1), dim3(block), 0, s, sums, masses, (uint32_t)d,
                       (uint32_t)k, centroids, old_centroids, max_shift);
    CLK_HIP_CHECK(hipGetLastError());
    return CLK_OK;
}
