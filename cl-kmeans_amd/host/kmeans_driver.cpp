#include "kmeans_driver.hpp"

#include <cstring>
#include <stdexcept>

#include "clkmeans_strategy.h"

// Sets the message clk_last_error() returns (defined with the C ABI).
extern "C" void clk_set_error_message(const char *msg);

namespace Clustering {

bool conf_is_double(const ConfigurationParser &cfg)
{
    const KmeansConfiguration km = cfg.get_kmeans_configuration();
    if (km.point_type == "double" && km.label_type == "uint64" && km.mass_type == "uint64")
        return true;
    if (km.point_type == "float" && km.label_type == "uint32" && km.mass_type == "uint32")
        return false;
    throw std::invalid_argument("Invalid type");  // bench.cpp:501-503
}

template <typename P, typename L, typename M>
typename ClusteringBenchmark<P, L, M>::ClClusteringFunction
make_kmeans(const ConfigurationParser &cfg, const DriverOptions &opt, std::shared_ptr<size_t> iterations_run)
{
    const KmeansConfiguration km = cfg.get_kmeans_configuration();
    if (conf_is_double(cfg) != std::is_same<P, double>::value)
        throw std::invalid_argument("Invalid type");
    ContextHandle ctx(opt.device);
    Queue queue(opt.device);
    std::shared_ptr<AbstractKmeans<P, L, M>> strategy;
    const std::string &pipe = km.pipeline;
    if (pipe == "single_stage") {
        auto s = std::make_shared<KmeansSingleStage<P, L, M>>();
        s->set_context(ctx);
        s->set_queue(queue);
        s->set_fused(cfg.get_fused_configuration());
        strategy = s;
    } else if (pipe == "single_stage_buffered") {
        auto s = std::make_shared<KmeansSingleStageBuffered<P, L, M>>();
        s->set_context(ctx);
        s->set_queue(queue);
        s->set_buffer_size(opt.buffer_size);
        s->set_fused(cfg.get_fused_configuration());
        strategy = s;
    } else if (pipe == "three_stage") {
        auto s = std::make_shared<KmeansThreeStage<P, L, M>>();
        s->set_context(ctx);
        s->set_queue(queue);
        s->set_labeler(cfg.get_labeling_configuration());
        s->set_mass_updater(cfg.get_mass_update_configuration());
        s->set_centroid_updater(cfg.get_centroid_update_configuration());
        strategy = s;
    } else if (pipe == "three_stage_buffered") {
        auto s = std::make_shared<KmeansThreeStageBuffered<P, L, M>>();
        s->set_context(ctx);
        s->set_queue(queue);
        s->set_buffer_size(opt.buffer_size);
        s->set_labeler(cfg.get_labeling_configuration());
        s->set_mass_updater(cfg.get_mass_update_configuration());
        s->set_centroid_updater(cfg.get_centroid_update_configuration());
        strategy = s;
    } else {
        throw std::invalid_argument(pipe);
    }
    strategy->set_convergence_tolerance(opt.tolerance);
    strategy->set_measurement(std::make_shared<Measurement::Measurement>());
    return [strategy, iterations_run](size_t iters, size_t d, std::shared_ptr<const std::vector<P>> pts,
                                      std::shared_ptr<std::vector<P>> cen, std::shared_ptr<std::vector<M>> ms,
                                      std::shared_ptr<std::vector<L>> lb) {
        auto m = (*strategy)(iters, d, pts, cen, ms, lb);
        if (iterations_run)
            *iterations_run = strategy->iterations_run();
        return m;
    };
}

template ClusteringBenchmarkF32::ClClusteringFunction make_kmeans<float, uint32_t, uint32_t>(
    const ConfigurationParser &, const DriverOptions &, std::shared_ptr<size_t>);
template ClusteringBenchmarkF64::ClClusteringFunction make_kmeans<double, uint64_t, uint64_t>(
    const ConfigurationParser &, const DriverOptions &, std::shared_ptr<size_t>);

// One clustering run of the .conf's strategy over host arrays of one triple.
template <typename P, typename L, typename M>
int kmeans_run(const char *conf_text, int device, uint64_t d, uint64_t n, uint64_t k, const P *host_points,
               P *host_centroids, M *host_masses, L *host_labels, uint64_t buffer_size, double tolerance,
               uint64_t *iterations_run, uint64_t *total_time_ns)
{
    try {
        if (!conf_text || !host_points || !host_centroids || !host_masses || !host_labels || d == 0 || k == 0)
            throw std::invalid_argument("clk_kmeans_run: null buffer or zero d/k");
        ConfigurationParser cfg;
        cfg.parse_string(conf_text);
        const KmeansConfiguration km = cfg.get_kmeans_configuration();
        if (cfg.has("kmeans.clusters") && km.clusters != k)
            throw std::invalid_argument("clk_kmeans_run: k differs from kmeans.clusters");
        DriverOptions opt;
        opt.device = device;
        if (buffer_size)
            opt.buffer_size = buffer_size;
        opt.tolerance = tolerance;
        auto iters = std::make_shared<size_t>(0);
        auto f = make_kmeans<P, L, M>(cfg, opt, iters);
        auto pts = std::make_shared<const std::vector<P>>(host_points, host_points + n * d);
        auto cen = std::make_shared<std::vector<P>>(host_centroids, host_centroids + k * d);
        auto ms = std::make_shared<std::vector<M>>(k);
        auto lb = std::make_shared<std::vector<L>>(n);
        auto m = f(km.iterations, d, pts, cen, ms, lb);
        std::memcpy(host_centroids, cen->data(), k * d * sizeof(P));
        std::memcpy(host_masses, ms->data(), k * sizeof(M));
        if (n)
            std::memcpy(host_labels, lb->data(), n * sizeof(L));
        if (iterations_run)
            *iterations_run = *iters;
        if (total_time_ns)
            *total_time_ns = m->total_by_name("TotalTime");
        return CLK_OK;
    } catch (const std::invalid_argument &e) {
        clk_set_error_message(e.what());
        return CLK_ERR_INVALID_ARGUMENT;
    } catch (const std::length_error &e) {
        clk_set_error_message(e.what());
        return CLK_ERR_SIZE;
    } catch (const std::exception &e) {
        clk_set_error_message(e.what());
        return CLK_ERR_HIP;
    }
}

}  // namespace Clustering

extern "C" int clk_kmeans_run(const char *conf_text, int device, uint64_t d, uint64_t n, uint64_t k,
                              const float *host_points, float *host_centroids, uint32_t *host_masses,
                              uint32_t *host_labels, uint64_t buffer_size, double tolerance,
                              uint64_t *iterations_run, uint64_t *total_time_ns)
{
    return Clustering::kmeans_run<float, uint32_t, uint32_t>(conf_text, device, d, n, k, host_points,
                                                            host_centroids, host_masses, host_labels,
                                                            buffer_size, tolerance, iterations_run,
                                                            total_time_ns);
}

extern "C" int clk_kmeans_run_f64(const char *conf_text, int device, uint64_t d, uint64_t n, uint64_t k,
                                  const double *host_points, double *host_centroids, uint64_t *host_masses,
                                  uint64_t *host_labels, uint64_t buffer_size, double tolerance,
                                  uint64_t *iterations_run, uint64_t *total_time_ns)
{
    return Clustering::kmeans_run<double, uint64_t, uint64_t>(conf_text, device, d, n, k, host_points,
                                                             host_centroids, host_masses, host_labels,
                                                             buffer_size, tolerance, iterations_run,
                                                             total_time_ns);
}

extern "C" int clk_conf_describe(const char *conf_text, char *out, size_t out_len)
{
    using namespace Clustering;
    try {
        if (!conf_text || !out)
            throw std::invalid_argument("clk_conf_describe: null argument");
        ConfigurationParser cfg;
        cfg.parse_string(conf_text);
        std::string s;
        auto kv = [&](const std::string &k, const std::string &v) { s += k + "=" + v + "\n"; };
        auto sizes = [](const size_t v[3]) {
            return std::to_string(v[0]) + "," + std::to_string(v[1]) + "," + std::to_string(v[2]);
        };
        auto b = cfg.get_benchmark_configuration();
        kv("benchmark.runs", std::to_string(b.runs));
        kv("benchmark.verify", b.verify ? "true" : "false");
        auto km = cfg.get_kmeans_configuration();
        kv("kmeans.clusters", std::to_string(km.clusters));
        kv("kmeans.pipeline", km.pipeline);
        kv("kmeans.iterations", std::to_string(km.iterations));
        kv("kmeans.converge", km.converge ? "true" : "false");
        kv("kmeans.types.point", km.point_type);
        kv("kmeans.types.label", km.label_type);
        kv("kmeans.types.mass", km.mass_type);
        auto l = cfg.get_labeling_configuration();
        kv("kmeans.labeling.strategy", l.strategy);
        kv("kmeans.labeling.global_size", sizes(l.global_size));
        kv("kmeans.labeling.local_size", sizes(l.local_size));
        kv("kmeans.labeling.vector_length", std::to_string(l.vector_length));
        kv("kmeans.labeling.unroll_clusters_length", std::to_string(l.unroll_clusters_length));
        kv("kmeans.labeling.unroll_features_length", std::to_string(l.unroll_features_length));
        auto m = cfg.get_mass_update_configuration();
        kv("kmeans.mass_update.strategy", m.strategy);
        kv("kmeans.mass_update.global_size", sizes(m.global_size));
        kv("kmeans.mass_update.local_size", sizes(m.local_size));
        kv("kmeans.mass_update.vector_length", std::to_string(m.vector_length));
        auto c = cfg.get_centroid_update_configuration();
        kv("kmeans.centroid_update.strategy", c.strategy);
        kv("kmeans.centroid_update.global_size", sizes(c.global_size));
        kv("kmeans.centroid_update.local_size", sizes(c.local_size));
        kv("kmeans.centroid_update.local_features", std::to_string(c.local_features));
        kv("kmeans.centroid_update.thread_features", std::to_string(c.thread_features));
        kv("kmeans.centroid_update.vector_length", std::to_string(c.vector_length));
        auto f = cfg.get_fused_configuration();
        kv("kmeans.fused.strategy", f.strategy);
        kv("kmeans.fused.global_size", sizes(f.global_size));
        kv("kmeans.fused.local_size", sizes(f.local_size));
        kv("kmeans.fused.vector_length", std::to_string(f.vector_length));
        kv("kmeans.fused.platform", std::to_string(f.platform));
        kv("kmeans.fused.device", std::to_string(f.device));
        if (s.size() + 1 > out_len)
            throw std::length_error("clk_conf_describe: output buffer too small");
        std::memcpy(out, s.c_str(), s.size() + 1);
        return CLK_OK;
    } catch (const std::invalid_argument &e) {
        clk_set_error_message(e.what());
        return CLK_ERR_INVALID_ARGUMENT;
    } catch (const std::length_error &e) {
        clk_set_error_message(e.what());
        return CLK_ERR_SIZE;
    } catch (const std::exception &e) {
        clk_set_error_message(e.what());
        return CLK_ERR_HIP;
    }
}
