// The .conf tuning surface: the reference's configuration structs
// (benchmark_configuration.hpp:17-20, kmeans_configuration.hpp:18-26,
// labeling_configuration.hpp:18-27, mass_update_configuration.hpp:18-25,
// centroid_update_configuration.hpp:18-27, fused_configuration.hpp:18-25)
// and ConfigurationParser (configuration_parser.cpp:23-328) with the same
// keys.  A size of 0 means "chosen for the GPU".
#pragma once

#include <cstddef>
#include <map>
#include <string>
#include <vector>

namespace Clustering {

struct BenchmarkConfiguration {
    size_t runs = 1;
    bool verify = false;
};

struct KmeansConfiguration {
    size_t clusters = 0;
    std::string pipeline;
    size_t iterations = 0;
    bool converge = false;
    std::string point_type = "float";
    std::string label_type = "uint32";
    std::string mass_type = "uint32";
};

struct LabelingConfiguration {
    size_t platform = 0, device = 0;
    std::string strategy = "unroll_vector";
    size_t global_size[3] = {0, 1, 1};
    size_t local_size[3] = {0, 1, 1};
    size_t vector_length = 4;
    size_t unroll_clusters_length = 1;
    size_t unroll_features_length = 1;
};

struct MassUpdateConfiguration {
    size_t platform = 0, device = 0;
    std::string strategy = "part_private";
    size_t global_size[3] = {0, 1, 1};
    size_t local_size[3] = {0, 1, 1};
    size_t vector_length = 4;
};

struct CentroidUpdateConfiguration {
    size_t platform = 0, device = 0;
    std::string strategy = "cluster_merge";
    size_t global_size[3] = {0, 1, 1};
    size_t local_size[3] = {0, 1, 1};
    size_t local_features = 1;
    size_t thread_features = 1;
    size_t vector_length = 4;
};

struct FusedConfiguration {
    size_t platform = 0, device = 0;
    std::string strategy = "cluster_merge";
    size_t global_size[3] = {0, 1, 1};
    size_t local_size[3] = {0, 1, 1};
    size_t vector_length = 4;
};

// INI-style key = value file with [section] headers, '#' comments; the full
// key is section + "." + key (boost::program_options config-file rules).
// Unknown keys, malformed values and more than 3 sizes throw
// std::invalid_argument; a repeated size key appends a dimension.
class ConfigurationParser {
public:
    void parse_file(const std::string &file);
    void parse_string(const std::string &text);

    BenchmarkConfiguration get_benchmark_configuration() const;
    KmeansConfiguration get_kmeans_configuration() const;
    LabelingConfiguration get_labeling_configuration() const;
    MassUpdateConfiguration get_mass_update_configuration() const;
    CentroidUpdateConfiguration get_centroid_update_configuration() const;
    FusedConfiguration get_fused_configuration() const;

    bool has(const std::string &key) const { return values_.count(key) != 0; }

private:
    const std::vector<std::string> *find(const std::string &key) const;
    std::map<std::string, std::vector<std::string>> values_;
};

}  // namespace Clustering
