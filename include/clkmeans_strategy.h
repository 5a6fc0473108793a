/*
 * clkmeans_strategy.h -- C ABI of the strategy level: one whole clustering
 * run driven by a .conf text, as the reference's bench runs one
 * AbstractKmeans::operator() (abstract_kmeans.hpp:106-130) built from
 * ConfigurationParser (configuration_parser.cpp:23-328) and the pipeline
 * switch of bench.cpp:227-390.  Host buffers in and out.
 */
#ifndef CLKMEANS_STRATEGY_H
#define CLKMEANS_STRATEGY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * conf_text: contents of a .conf file ([kmeans] pipeline = single_stage |
 * three_stage | single_stage_buffered | three_stage_buffered, iterations,
 * per-stage strategy and sizes).  host_points is float [d][n];
 * host_centroids is float [d][k], initial centroids in, final out;
 * host_masses (k) and host_labels (n) are out.  buffer_size: bytes of points
 * per chunk for the _buffered pipelines (0 = 256 MiB).  tolerance > 0 stops
 * once the largest centroid move is <= tolerance.  iterations_run and
 * total_time_ns (the reference's TotalTime, host clock around the loop) may
 * be NULL.  Unknown keys / strategy names: CLK_ERR_INVALID_ARGUMENT with the
 * message in clk_last_error().
 */
int clk_kmeans_run(const char *conf_text, int device, uint64_t d, uint64_t n, uint64_t k,
                   const float *host_points, float *host_centroids, uint32_t *host_masses,
                   uint32_t *host_labels, uint64_t buffer_size, double tolerance,
                   uint64_t *iterations_run, uint64_t *total_time_ns);

/*
 * The same for a .conf whose [kmeans] types are point = double, label =
 * uint64, mass = uint64 (bench.cpp:469-481): double points and centroids,
 * uint64 masses and labels.  clk_kmeans_run takes the float / uint32 /
 * uint32 types; a .conf naming the other triple is rejected with
 * CLK_ERR_INVALID_ARGUMENT ("Invalid type", bench.cpp:501-503).
 */
int clk_kmeans_run_f64(const char *conf_text, int device, uint64_t d, uint64_t n, uint64_t k,
                       const double *host_points, double *host_centroids, uint64_t *host_masses,
                       uint64_t *host_labels, uint64_t buffer_size, double tolerance,
                       uint64_t *iterations_run, uint64_t *total_time_ns);

/*
 * Parse conf_text with the ConfigurationParser rules and write the resulting
 * configuration structs as "key=value" lines (every field, defaults
 * included) into out.  No GPU is touched.  CLK_ERR_INVALID_ARGUMENT for an
 * unknown key or malformed value, CLK_ERR_SIZE if out_len is too small.
 */
int clk_conf_describe(const char *conf_text, char *out, size_t out_len);

#ifdef __cplusplus
}
#endif

#endif /* CLKMEANS_STRATEGY_H */
