// Builds the strategy a .conf names (bench.cpp:227-390) as a clustering
// function over host vectors.
#pragma once

#include <memory>
#include <string>

#include "benchmark.hpp"
#include "configuration.hpp"
#include "kmeans.hpp"

namespace Clustering {

struct DriverOptions {
    int device = 0;
    size_t buffer_size = 256ull << 20;
    double tolerance = 0.0;
};

// The returned function keeps the strategy alive; iterations_run (if given)
// receives the number of iterations of the last call.
ClusteringBenchmark::ClClusteringFunction make_kmeans(const ConfigurationParser &cfg,
                                                      const DriverOptions &opt,
                                                      std::shared_ptr<size_t> iterations_run = nullptr);

}  // namespace Clustering
