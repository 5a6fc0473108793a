/*
 * clkmeans.h -- C ABI of the MI355X (gfx950) Lloyd-iteration library.
 *
 * Every compute entry point replaces one operator of CL-kmeans (the
 * std::function objects its factories return) and keeps that operator's
 * argument meaning: sizes first, then device buffers, accumulate-vs-overwrite
 * semantics unchanged.  Reference citations are file:line in the CL-kmeans
 * source tree.
 *
 * Conventions
 *  - All pointers named points/centroids/labels/masses/sums/data are DEVICE
 *    pointers (hipMalloc memory) unless the name starts with host_.
 *  - Layouts are the reference's column-major ones:
 *      points     float[d][n]   points[f * n + p]    (binary_format.cpp:37-43)
 *      centroids  float[d][k]   centroids[f * k + c] (lloyd_fused_feature_sum.cl:230-232)
 *      sums       double[d][k]  same layout as centroids (FP64 accumulator)
 *      labels     uint32[n], masses uint32[k]
 *  - stream is a hipStream_t passed as void*; NULL is the default stream.
 *    Every compute call only enqueues work on that stream.
 *  - Return value: CLK_OK (0) or a negative CLK_ERR_*; clk_last_error()
 *    returns the calling thread's last message.  Unknown strategy names give
 *    CLK_ERR_INVALID_ARGUMENT (the reference throws std::invalid_argument,
 *    fused_factory.hpp:84); size mismatches give CLK_ERR_SIZE (the reference
 *    asserts, fused_feature_sum.hpp:211-220).
 *  - A context owns scratch (per-block partials) and is used by one host
 *    thread at a time, like one reference operator object.
 *
 * Accumulator type.  The reference accumulates centroid sums in PointT
 * (float).  This library accumulates them exactly (integers on a
 * per-feature grid, see "exact sums" below), hands them over as FP64
 * (double[d][k]) and divides in FP64, rounding once to float; masses stay
 * uint32.
 */
#ifndef CLKMEANS_H
#define CLKMEANS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLK_OK 0
#define CLK_ERR_INVALID_ARGUMENT (-1)
#define CLK_ERR_SIZE (-2)
#define CLK_ERR_HIP (-3)
#define CLK_ERR_UNSUPPORTED (-4)
#define CLK_ERR_NOMEM (-5)

typedef struct clk_context clk_context;

/* Stage ids for clk_strategy_id / clk_kernel_config.stage. */
#define CLK_STAGE_FUSED 0
#define CLK_STAGE_LABELING 1
#define CLK_STAGE_MASS_UPDATE 2
#define CLK_STAGE_CENTROID_UPDATE 3

/* Strategy ids (the factories' strategy strings). */
#define CLK_FUSED_CLUSTER_MERGE 0   /* "cluster_merge"  fused_factory.hpp:73-77 */
#define CLK_FUSED_FEATURE_SUM 1     /* "feature_sum"    fused_factory.hpp:78-82 */
/* Extension: "cluster_merge" that reads the labels buffer first.  A point
 * keeps its previous label, with one distance computed, when that distance
 * is inside the label's keep radius -- (1-2^-10)/(2+2^-10)^2 of the squared
 * distance to the nearest other centroid, so the triangle inequality plus
 * the FP32 rounding bound prove the full scan returns the same label (the
 * reference's did-changes test, kmeans_naive.cpp:65-68, stores only changed
 * labels the same way).  Other points scan every centroid.  Labels, masses
 * and sums are bit-identical to "cluster_merge" for any labels content. */
#define CLK_FUSED_CLUSTER_MERGE_BOUNDED 2 /* "cluster_merge_bounded" */
#define CLK_LABELING_UNROLL_VECTOR 0 /* "unroll_vector" labeling_factory.hpp:63-80 */
/* Extensions (no reference counterpart; the factory would throw for them):
 * x·c^T on MFMA with 16-bit operands and FP32 accumulation, then the exact
 * FP32 distance for every point whose best-to-second margin is inside the
 * derived error bound -- the labels equal "unroll_vector"'s.  Used for
 * d <= 128 and n < 2^32 when the 16-bit centroid table fits in LDS; other
 * sizes run the "unroll_vector" kernel. */
#define CLK_LABELING_MFMA_BF16 1    /* "mfma_bf16" */
#define CLK_LABELING_MFMA_FP16 2    /* "mfma_fp16" */
/* Extension: "unroll_vector" that reads the labels buffer first and keeps a
 * point's previous label after one distance where the keep radius of
 * CLK_FUSED_CLUSTER_MERGE_BOUNDED proves the scan would return it; only
 * changed labels are stored.  The labels equal "unroll_vector"'s. */
#define CLK_LABELING_UNROLL_VECTOR_BOUNDED 3 /* "unroll_vector_bounded" */
#define CLK_MASS_GLOBAL_ATOMIC 0    /* "global_atomic"  mass_update_factory.hpp:71-75 */
#define CLK_MASS_PART_GLOBAL 1      /* "part_global"    mass_update_factory.hpp:76-80 */
#define CLK_MASS_PART_LOCAL 2       /* "part_local"     mass_update_factory.hpp:81-85 */
#define CLK_MASS_PART_PRIVATE 3     /* "part_private"   mass_update_factory.hpp:86-90 */
#define CLK_CENTROID_FEATURE_SUM 0        /* "feature_sum"        centroid_update_factory.hpp:83-92 */
#define CLK_CENTROID_FEATURE_SUM_PARDIM 1 /* "feature_sum_pardim" centroid_update_factory.hpp:93-111 */
#define CLK_CENTROID_CLUSTER_MERGE 2      /* "cluster_merge"      centroid_update_factory.hpp:112-120 */

/*
 * Launch geometry and strategy of one stage: the reference's
 * {Fused,Labeling,MassUpdate,CentroidUpdate}Configuration structs
 * (fused_configuration.hpp:18-25, labeling_configuration.hpp:18-27,
 * mass_update_configuration.hpp:18-25, centroid_update_configuration.hpp:18-27)
 * minus platform/device.  A size of 0 means "choose for this GPU".
 * global_size counts work-items (threads), as in the reference.
 */
typedef struct clk_kernel_config {
    int32_t stage;
    int32_t strategy;
    uint64_t global_size[3];
    uint64_t local_size[3];
    uint64_t vector_length;   /* points per work-item per load: 1, 2, 4 or 8 */
    uint64_t local_features;  /* feature_sum_pardim: work-items per block along features */
    uint64_t thread_features; /* feature_sum_pardim: features per work-item */
} clk_kernel_config;

/* ---- errors, context ------------------------------------------------- */

const char *clk_last_error(void);

/* Create a context on HIP device `device` (scratch, device properties). */
int clk_context_create(int device, clk_context **out);
int clk_context_destroy(clk_context *ctx);

/* Map a factory strategy string to its id; CLK_ERR_INVALID_ARGUMENT when the
 * reference factory would throw std::invalid_argument (fused_factory.hpp:84,
 * labeling_factory.hpp:83, mass_update_factory.hpp:92,
 * centroid_update_factory.hpp:122). */
int clk_strategy_id(int stage, const char *name);

/* Fill cfg with this library's defaults for the stage/strategy. */
int clk_default_config(clk_context *ctx, int stage, int strategy, clk_kernel_config *cfg);

/* ---- per-op timing (the reference's Measurement DataPoints) ------------ */

/* When enabled, every kernel an op enqueues is bracketed by HIP events
 * (measurement.cpp:37-60 sums OpenCL event durations the same way). */
int clk_profiling_enable(clk_context *ctx, int enable);
/* Tag subsequent records with an iteration number (-1 = none). */
int clk_profiling_set_iteration(clk_context *ctx, int iteration);
int clk_profiling_clear(clk_context *ctx);
/* Number of kernel records so far. */
int64_t clk_profiling_count(clk_context *ctx);
/* Record i: kernel name, iteration, elapsed milliseconds (waits for it). */
int clk_profiling_get(clk_context *ctx, int64_t i, char *name, size_t name_len,
                      int32_t *iteration, float *ms);

/* ---- the Lloyd operators ------------------------------------------------ */

/*
 * Fused single pass.  Replaces FusedFactory::FusedFunction
 * (fused_factory.hpp:34-53) as implemented by FusedClusterMerge
 * (fused_cluster_merge.hpp:193-414) and FusedFeatureSum
 * (fused_feature_sum.hpp:191-444):
 *   labels[p]  = argmin_c sum_f fma-sequence (x_f - c_f)^2   (overwrite)
 *   masses[c] += #{p : labels[p] == c}                       (accumulate)
 *   sums[f,c] += sum_{p : labels[p] == c} x_f                (accumulate)
 * old_centroids is read only; sums / masses must not alias it.  One pass
 * over the points when k*d fits the work-group LDS tiles; otherwise the
 * labeling, mass and centroid stages run one after another (the same
 * results).
 */
int clk_fused(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
              uint64_t n, uint64_t k, const float *points,
              const float *old_centroids, double *sums, uint32_t *labels,
              uint32_t *masses, void *stream);

/*
 * One whole Lloyd iteration of the single_stage pipeline: the body of
 * KmeansSingleStage::run's loop (kmeans_single_stage.hpp:92-151: zero
 * masses and sums, fused pass, divide), as one kernel launch on the LDS-tile
 * paths.  labels, masses and sums are OVERWRITTEN (not accumulated);
 * new_centroids[f,c] = (float)(sums[f,c] / (double)masses[c]) (NaN for an
 * empty cluster); when max_shift (device double) is non-NULL it receives
 * max_c ||new_c - old_c||_2 over non-empty clusters.  new_centroids must not
 * alias old_centroids.  Calls sharing a context must be ordered on one stream.
 */
int clk_lloyd_step(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                   uint64_t n, uint64_t k, const float *points,
                   const float *old_centroids, float *new_centroids, double *sums,
                   uint32_t *labels, uint32_t *masses, double *max_shift, void *stream);

/*
 * No counterpart in the reference: clk_lloyd_step for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and stream
 * semantics as clk_lloyd_step; it calls no hipMalloc / hipFree.  Its scratch
 * (per-work-group partial sums and masses, and the sum grid of the call's
 * points unless one is pinned) lives in one static device array of
 * CLK_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; the
 * persistent grid is clamped to the partial rows that fit.  CLK_ERR_SIZE when
 * one row, k*(8d+4) bytes, does not fit; CLK_ERR_UNSUPPORTED when k*d needs
 * the separate labeling / mass / centroid stages (beyond the LDS tiles).
 * Every call on one GPU shares the array and leaves its contents there:
 * calls that overlap in time on one GPU race (order them on one stream).
 * No profiling records are taken.
 */
#define CLK_NOALLOC_SCRATCH_BYTES (8u << 20)
int clk_lloyd_step_noalloc(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                           uint64_t n, uint64_t k, const float *points,
                           const float *old_centroids, float *new_centroids, double *sums,
                           uint32_t *labels, uint32_t *masses, double *max_shift, void *stream);

/*
 * Labeling only.  Replaces LabelingFactory::LabelingFunction
 * (labeling_factory.hpp:33-48) / LabelingUnrollVector
 * (labeling_unroll_vector.hpp:168-272): same arithmetic as clk_fused.
 */
int clk_labeling(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                 uint64_t n, uint64_t k, const float *points,
                 const float *centroids, uint32_t *labels, void *stream);

/* Diagnostic of the MFMA labeling strategies: number of points of the last
 * MFMA labeling call on this context whose margin test failed and that went
 * through the exact refinement (synchronizes the device; 0 before any MFMA
 * call).  No reference counterpart. */
int clk_labeling_refined_points(clk_context *ctx, uint64_t *points);

/* Keep a 16-bit copy of device points[d][n] in the operand order of the MFMA
 * labeling `strategy` (CLK_LABELING_MFMA_BF16 or CLK_LABELING_MFMA_FP16;
 * built on stream): 64 * ceil(d / 32) + 4 bytes per point.  Later MFMA
 * labelings of that strategy on this same buffer, n and d read the copy --
 * one contiguous stream -- instead of d strided FP32 rows per point (those
 * touch d pages per 64 points, more than the address-translation cache
 * holds); the exact refinement still reads the FP32 points, so the labels
 * are the same.  Pin again after changing the buffer's contents.  A new pin
 * replaces the old one.  No reference counterpart. */
int clk_labeling_pin_points(clk_context *ctx, int32_t strategy, uint64_t d, uint64_t n,
                            const float *points, void *stream);
/* Drop the copy (its memory is kept for the next pin until the context is
 * destroyed). */
int clk_labeling_unpin_points(clk_context *ctx);

/*
 * masses[c] += #{p : labels[p] == c}.  Replaces
 * MassUpdateFactory::MassUpdateFunction (mass_update_factory.hpp:36-48) and
 * MassUpdate{GlobalAtomic,PartGlobal,PartLocal,PartPrivate}
 * (mass_update_global_atomic.hpp:110-156, mass_update_part_global.hpp:104-189,
 * mass_update_part_local.hpp:125-217, mass_update_part_private.hpp:138-249).
 * Labels >= k are rejected by no kernel; callers pass labels from clk_labeling.
 */
int clk_mass_update(clk_context *ctx, const clk_kernel_config *cfg, uint64_t n,
                    uint64_t k, const uint32_t *labels, uint32_t *masses,
                    void *stream);

/*
 * sums[f,c] += sum_{p : labels[p] == c} x_f.  Replaces
 * CentroidUpdateFactory::CentroidUpdateFunction
 * (centroid_update_factory.hpp:37-54) and CentroidUpdate{FeatureSum,
 * FeatureSumPardim,ClusterMerge} (centroid_update_feature_sum.hpp:116-206,
 * centroid_update_feature_sum_pardim.hpp:149-306,
 * centroid_update_cluster_merge.hpp:145-272).  masses is accepted and unused,
 * as in the reference kernels.
 */
int clk_centroid_update(clk_context *ctx, const clk_kernel_config *cfg,
                        uint64_t d, uint64_t n, uint64_t k, const float *points,
                        double *sums, const uint32_t *labels,
                        const uint32_t *masses, void *stream);

/*
 * centroids[f,c] = (float)(sums[f,c] / (double)masses[c]); an empty cluster
 * gives NaN (0/0).  Replaces MatrixBinaryOp::row with Divide
 * (matrix_binary_op.hpp:87-127, matrix_binary_op.cl:23-39), computed in FP64.
 * When old_centroids and max_shift are both non-NULL, *max_shift (device
 * double) receives max_c ||new_c - old_c||_2 over non-empty clusters.
 */
int clk_matrix_divide_row(clk_context *ctx, uint64_t d, uint64_t k,
                          const double *sums, const uint32_t *masses,
                          float *centroids, const float *old_centroids,
                          double *max_shift, void *stream);

/* fst[i] += snd[i], i < count.  MatrixBinaryOp::matrix with Add
 * (matrix_binary_op.hpp:149-188, matrix_binary_op.cl:51-65). */
int clk_matrix_add_u32(clk_context *ctx, uint64_t count, uint32_t *fst,
                       const uint32_t *snd, void *stream);
int clk_matrix_add_f64(clk_context *ctx, uint64_t count, double *fst,
                       const double *snd, void *stream);

/*
 * In-place column-vector sum: data holds num_cols vectors of result_rows
 * elements (data[col * result_rows + row]); afterwards
 * data[row] = sum over col, for row < result_rows, in a fixed order
 * (deterministic).  Replaces ReduceVectorParcol<T>
 * (reduce_vector_parcol.hpp:30-167); any num_cols >= 1 is accepted.
 */
int clk_reduce_vector_parcol_u32(clk_context *ctx, uint64_t num_cols,
                                 uint64_t result_rows, uint32_t *data,
                                 void *stream);
int clk_reduce_vector_parcol_f64(clk_context *ctx, uint64_t num_cols,
                                 uint64_t result_rows, double *data,
                                 void *stream);

/* Zero `bytes` bytes at a device pointer with a kernel (usable under HIP graph
 * capture).  Replaces boost::compute::fill_async(..., 0, queue) used at
 * kmeans_single_stage.hpp:97-112. */
int clk_fill_zero(clk_context *ctx, void *ptr, uint64_t bytes, void *stream);

/* centroids[f,c] = points[f, c mod n] on the device: KmeansInitializer::first_x
 * (kmeans_initializer.cpp:29-39). */
int clk_init_first_x(clk_context *ctx, uint64_t d, uint64_t n, uint64_t k,
                     const float *points, float *centroids, void *stream);

/* ---- exact sums: the sum grid ----------------------------------------- */
/*
 * No reference counterpart: the reference's FP32 sums follow its kernel
 * geometry (per-work-item tiles, lloyd_fused_cluster_merge.cl, merged by
 * reduce_vector_parcol.cl's tree), so its centroid bits depend on the
 * launch.  Here every op that accumulates centroid sums (clk_fused,
 * clk_lloyd_step, clk_centroid_update and the batch engine) first rounds
 * each coordinate to nearest-even on a per-feature power-of-two grid and
 * adds the grid integers exactly: sums, centroids and therefore labels do
 * not depend on the strategy, the launch geometry, chunking, the batch tile
 * size or the other jobs of a batch.
 *   unit of feature f = 2^(E_f - b), where M_f = max |x_f| over the finite
 *   coordinates, E_f = the least integer with M_f < 2^E_f (at least -100)
 *   and b = min(25, 53 - ceil(log2(n + 1)));
 * a coordinate moves by at most half a unit (<= 2^-26 M_f).  A NaN / +-inf
 * coordinate makes its cell's sum NaN / +-inf (NaN for both infinities).
 * Without a pinned grid each op uses the grid of its own (points, n): one
 * extra read of the points per call.  A pinned grid serves every later op of
 * the context with the same d until cleared or replaced: the strategies
 * (clk_kmeans_run, the C++ classes) pin the grid of the whole data set once,
 * so their chunked (_buffered) pipelines give the unchunked sums.  A point
 * outside a pinned grid's range (|x_f| >= 2^E_f) makes its cell's sum NaN.
 */
/* Compute the grid of device points[d][n] (on stream) and pin it.  Later
 * ops on this same buffer and n take it as that data's own grid: when the
 * data hold no NaN or infinity their coordinates are quantized without a
 * range check, so pin again (or clear) after changing the buffer's
 * contents.  Ops on other buffers check every coordinate against it. */
int clk_sum_grid_from_points(clk_context *ctx, uint64_t d, uint64_t n,
                             const float *points, void *stream);
/* Pin the grid of n points whose per-feature max |x| the caller knows
 * (host_max_abs[d]; e.g. host-resident data streamed in chunks). */
int clk_sum_grid_from_max(clk_context *ctx, uint64_t d, uint64_t n,
                          const float *host_max_abs);
/* Unpin: later ops compute the grid of their own points again. */
int clk_sum_grid_clear(clk_context *ctx);
/* The units (host_units[d]) an op of this context would use for
 * (d, n, points): the pinned grid's, or the grid of these points.
 * Synchronizes the stream. */
int clk_sum_grid_units(clk_context *ctx, uint64_t d, uint64_t n,
                       const float *points, double *host_units, void *stream);

/* ---- the double / uint64 / uint64 type triple ------------------------- */
/*
 * The reference instantiates every strategy for double points with uint64
 * labels and masses as well (bench.cpp:469-503; the OpenCL kernels built
 * with CL_POINT=double, CL_POINT_MAX=DBL_MAX, CL_LABEL/CL_MASS=ulong,
 * fused_feature_sum.hpp:80-95, labeling_unroll_vector.hpp:64-76).  Same
 * layouts and semantics as the float entry points above with double points
 * and centroids, uint64 labels and masses, FP64 sums: the distance is the
 * sequential FP64 fma from 0, argmin with strict '<' from DBL_MAX, label 0
 * when nothing is smaller; every strategy id of a stage runs the same FP64
 * pass (an LDS-tiled pass with a fixed-order merge where the tile fits, FP64
 * / uint64 global atomics beyond).  Labels and masses are exact; sums carry
 * FP64 rounding of a different summation order.
 */
int clk_fused_f64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                  uint64_t n, uint64_t k, const double *points,
                  const double *old_centroids, double *sums, uint64_t *labels,
                  uint64_t *masses, void *stream);
int clk_lloyd_step_f64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                       uint64_t n, uint64_t k, const double *points,
                       const double *old_centroids, double *new_centroids,
                       double *sums, uint64_t *labels, uint64_t *masses,
                       double *max_shift, void *stream);
int clk_labeling_f64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t d,
                     uint64_t n, uint64_t k, const double *points,
                     const double *centroids, uint64_t *labels, void *stream);
int clk_mass_update_u64(clk_context *ctx, const clk_kernel_config *cfg, uint64_t n,
                        uint64_t k, const uint64_t *labels, uint64_t *masses,
                        void *stream);
int clk_centroid_update_f64(clk_context *ctx, const clk_kernel_config *cfg,
                            uint64_t d, uint64_t n, uint64_t k, const double *points,
                            double *sums, const uint64_t *labels,
                            const uint64_t *masses, void *stream);
/* centroids[f,c] = sums[f,c] / (double)masses[c] (0/0 = NaN); optional
 * max shift as clk_matrix_divide_row. */
int clk_matrix_divide_f64(clk_context *ctx, uint64_t d, uint64_t k,
                          const double *sums, const uint64_t *masses,
                          double *centroids, const double *old_centroids,
                          double *max_shift, void *stream);
int clk_matrix_add_u64(clk_context *ctx, uint64_t count, uint64_t *fst,
                       const uint64_t *snd, void *stream);

/* ---- host-side synthetic data ----------------------------------------- */

/*
 * Gaussian blobs, feature-major float32 host_points[d][n]: point p belongs to
 * blob p mod k (equal sizes, interleaved); blob centres uniform in
 * [-box, box]^d; isotropic noise sigma.  Deterministic in (seed, n, d, k):
 * a counter-based generator, independent of thread count.  Optional
 * host_centres[d][k] receives the blob centres.  Takes the role of the
 * reference's cluster_generator.cpp, which instead writes points grouped by
 * blob (cluster_generator.cpp:178-195); interleaving makes first_x pick one
 * seed from each blob.
 */
int clk_generate_blobs(uint64_t n, uint64_t d, uint64_t k, double sigma,
                       double box, uint64_t seed, float *host_points,
                       float *host_centres, int threads);

/* The generator CLI's form (replaces generator.cpp:36-53 with
 * cluster_generator.cpp:39-100): centres uniform in [domain_min,
 * domain_max]^d, noise N(0, radius) per feature (the reference draws
 * N(-radius, radius), cluster_generator.cpp:69); grouped != 0 puts the
 * points of a blob next to each other (blob of p = p / ceil(n/k)) as the
 * reference does, else interleaved as clk_generate_blobs. */
int clk_generate_points(uint64_t n, uint64_t d, uint64_t k, double radius,
                        double domain_min, double domain_max, uint64_t seed,
                        int grouped, float *host_points, float *host_centres,
                        int threads);

#ifdef __cplusplus
}
#endif

#endif /* CLKMEANS_H */
