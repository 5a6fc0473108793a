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
// receives the number of iterations of the last call.  The .conf's types
// must be the template's triple (std::invalid_argument("Invalid type")
// otherwise, bench.cpp:502).
template <typename PointT, typename LabelT, typename MassT>
typename ClusteringBenchmark<PointT, LabelT, MassT>::ClClusteringFunction
make_kmeans(const ConfigurationParser &cfg, const DriverOptions &opt,
            std::shared_ptr<size_t> iterations_run = nullptr);

// Whether the .conf names the double/uint64/uint64 triple (else it must name
// float/uint32/uint32).
bool conf_is_double(const ConfigurationParser &cfg);

}  // namespace Clustering
