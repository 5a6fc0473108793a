/*
 * clkmeans_batch.h -- many independent k-means jobs on one GPU.
 *
 * No reference counterpart: CL-kmeans runs one clustering per process
 * (bench.cpp:200-452 drives one AbstractKmeans at a time).  A batch keeps many
 * jobs (uneven n, d, k) resident in device memory and advances all active jobs
 * by one Lloyd iteration per persistent kernel launch, captured once as a HIP
 * graph.  Every job follows the reference's iteration exactly as
 * clk_lloyd_step does (labels: the sequential-fma distance with strict '<'
 * and lowest index on ties; masses exact; FP64 sums and divide), and a job's
 * result does not depend on the other jobs of the batch: its points are cut
 * into fixed tiles, each tile's partial sums go to its own FP64 row, and the
 * work-group that finishes a job's last tile merges the rows in tile order.
 *
 * Layouts as in clkmeans.h: points float[d][n], centroids float[d][k], labels
 * and masses uint32.  Jobs are addressed by slot; a slot is free again after
 * clk_batch_take.  Functions return CLK_OK or a negative CLK_ERR_* code.
 */
#ifndef CLKMEANS_BATCH_H
#define CLKMEANS_BATCH_H

#include "clkmeans.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clk_batch clk_batch;

/* Job states. */
#define CLK_JOB_FREE 0
#define CLK_JOB_ACTIVE 1
#define CLK_JOB_DONE 2

typedef struct clk_batch_config {
    uint32_t max_jobs;        /* slots */
    uint64_t arena_bytes;     /* device memory for job data (points, labels, centroids, partials) */
    uint32_t tile_work;       /* points per tile per cluster (tiles of 1024..16384 points); 0 = 128;
                                 clk_batch_autotune picks one */
    uint32_t max_d;           /* largest d admitted (<= 64) */
    uint32_t max_k;           /* largest k admitted */
    int32_t labeling;         /* CLK_LABELING_UNROLL_VECTOR (plain FP32), CLK_LABELING_MFMA_BF16 or
                                 CLK_LABELING_MFMA_FP16: x.c^T on the matrix cores with the exact
                                 refinement (same labels) */
} clk_batch_config;

/* Device memory one job needs in the arena (for admission by footprint). */
uint64_t clk_batch_job_bytes(const clk_batch_config *cfg, uint64_t n, uint32_t d, uint32_t k);

int clk_batch_create(clk_context *ctx, const clk_batch_config *cfg, clk_batch **out);
int clk_batch_destroy(clk_batch *b);

/* Admit a job: copy host points float[d][n] and initial centroids float[d][k]
 * to the device.  It runs until max_iterations iterations or, when
 * tolerance > 0, until the max centroid shift (FP64, ||new_c - old_c||_2 over
 * non-empty clusters) falls below tolerance.  CLK_ERR_NOMEM when no slot or
 * arena space is free (the caller retries after taking finished jobs).
 * CLK_ERR_SIZE unless n >= 1 and n*d <= 2^30 (the tile kernels address a
 * job's points with 32-bit byte offsets; larger data sets run through
 * clk_lloyd_step). */
int clk_batch_add(clk_batch *b, uint64_t n, uint32_t d, uint32_t k, const float *host_points,
                  const float *host_centroids, uint32_t max_iterations, double tolerance,
                  int32_t *slot, void *stream);

/* One Lloyd iteration of every active job: one persistent kernel launch
 * (replayed from a captured HIP graph while the set of jobs is unchanged). */
int clk_batch_step(clk_batch *b, void *stream);

/* Number of active jobs and, in states[] (max_jobs entries, may be NULL),
 * each slot's CLK_JOB_* state.  Synchronizes the stream. */
int clk_batch_poll(clk_batch *b, uint32_t *states, uint32_t *active, void *stream);

/* Copy a job's results to the host (any pointer may be NULL) and free its
 * slot.  centroids float[d][k], labels uint32[n] (of the last labeling),
 * masses uint32[k]; iterations run and the last max shift. */
int clk_batch_take(clk_batch *b, int32_t slot, float *host_centroids, uint32_t *host_labels,
                   uint32_t *host_masses, uint32_t *iterations, double *shift, void *stream);

/* Device buffers of a slot (points [d][n], current centroids [d][k], labels
 * [n]), e.g. to run the same job through clk_lloyd_step for comparison. */
int clk_batch_job_buffers(clk_batch *b, int32_t slot, const float **points, float **centroids,
                          uint32_t **labels);

/* The plain per-job loop on the same device buffers, for comparison: every
 * active job in slot order runs its remaining iterations alone through
 * clk_lloyd_step (fused cluster_merge pass + FP64 merge and divide), one job
 * after another; with a tolerance the shift is read back after every
 * iteration.  Jobs end in the same states as with clk_batch_step. */
int clk_batch_run_serial(clk_batch *b, void *stream);

/* Autotune tile_work on the current jobs: times label + accumulate passes
 * (no job advances) for tile_work in {4, 8, 16, 32, 64, 128, 256}, re-tiles
 * the jobs with the fastest and keeps it for jobs admitted later (the tuned
 * value takes the place of a hand-edited .conf entry).  Centroid sums are
 * exact, so the tiling changes no job's result and the call may come at any
 * point; after the jobs' first iteration the timed passes include the keep
 * test that every later iteration runs (the first iteration's full scans
 * favour smaller tiles).  chosen may be NULL.  CLK_ERR_INVALID_ARGUMENT if
 * no job is active. */
int clk_batch_autotune(clk_batch *b, uint32_t *chosen, void *stream);

/* ---- serving: jobs arriving over time --------------------------------- */

typedef struct clk_serve_job {
    uint64_t n;
    uint32_t d, k;
    const float *points;      /* host [d][n]; pinned memory lets the copy overlap the GPU work.
                                 With device_points set: device memory, read in place (it must
                                 stay valid until the job has finished) */
    int32_t device_points;
    const float *centroids;   /* host [d][k], initial centroids */
    double arrival;           /* seconds after the start of the run (non-decreasing over the array) */
    uint32_t max_iterations;
    double tolerance;         /* stop when the max centroid shift < tolerance (0: never) */
    float *out_centroids;     /* [d][k] or NULL */
    uint32_t *out_labels;     /* [n] or NULL */
    uint32_t *out_masses;     /* [k] or NULL */
    uint32_t iterations;      /* out */
    double shift;             /* out: the last max shift (0 when no tolerance) */
    double admitted;          /* out: seconds after the start when the copy-in began */
    double finished;          /* out: seconds after the start when the result was on the host */
} clk_serve_job;

typedef struct clk_serve_stats {
    double makespan;          /* first arrival .. last completion, seconds */
    double solves_per_s;      /* jobs / makespan */
    double p50, p99;          /* latency (finished - arrival), nearest rank, seconds */
    double mean_latency;
    uint64_t steps;           /* iterations launched (batch steps + lone-job steps) */
    uint32_t max_active;      /* most jobs iterating at once */
    uint64_t solo_steps;      /* of those, iterations of a lone job through the fused pass */
} clk_serve_stats;

/* Continuous batching over an empty batch: a job is admitted once its
 * arrival time has passed and a slot and arena space are free (in arrival
 * order), its points are copied in on a second stream while the resident jobs
 * iterate, it joins the next iteration's launch, and it leaves (results
 * copied to its out_* buffers) when it reaches its tolerance or iteration
 * limit.  While one job of at most 2^20 points is resident and no other has
 * arrived, its iterations run through the per-job fused pass (as
 * clk_lloyd_step, with the job's sum grid) instead of the batch step.  A
 * job's results are those it gets alone in a batch, whichever path ran each
 * of its iterations. */
int clk_serve_run(clk_batch *b, clk_serve_job *jobs, uint32_t count, clk_serve_stats *stats, void *stream);

/* The plain per-job loop for the same arrivals: one job at a time, copied in
 * and iterated with clk_lloyd_step (fused cluster_merge, FP32 labels, FP64
 * merge), the shift read back after every iteration when it has a
 * tolerance. */
int clk_serve_serial(clk_context *ctx, clk_serve_job *jobs, uint32_t count, clk_serve_stats *stats, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CLKMEANS_BATCH_H */
