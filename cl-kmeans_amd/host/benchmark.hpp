// The bench harness around the strategies: BinaryFormat (binary_format.cpp:20-46),
// KmeansInitializer (kmeans_initializer.cpp:16-39), KmeansNaive -- the CPU
// Lloyd that bench --verify compares against (kmeans_naive.cpp:31-106) --,
// ClusteringBenchmark / ClusteringBenchmarkStats (clustering_benchmark.cpp:135-374)
// and the mean squared error of Utility::mse (utility.hpp:29-51).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "measurement.hpp"

namespace Clustering {

// Column-major host matrix: rows = points (or clusters), cols = features,
// element (r, c) at data[c * rows + r] (matrix.hpp, ColMajor = true).
template <typename T>
struct Matrix {
    std::vector<T> data;
    size_t rows = 0, cols = 0;
    void resize(size_t r, size_t c)
    {
        rows = r;
        cols = c;
        data.assign(r * c, T());
    }
    T &operator()(size_t r, size_t c) { return data[c * rows + r]; }
    const T &operator()(size_t r, size_t c) const { return data[c * rows + r]; }
};

// .bin files: uint64 num_features, uint64 num_clusters (must be 0),
// uint64 num_points, then float32 [num_features][num_points].
struct BinaryFormat {
    static void read(const std::string &file, Matrix<float> &points);
    static void write(const std::string &file, const Matrix<float> &points);
};

struct KmeansInitializer {
    // centroid c = point (c mod n)
    static void first_x(const Matrix<float> &points, Matrix<float> &centroids);
    // centroid c = a uniformly random point (std::random_device seeded)
    static void forgy(const Matrix<float> &points, Matrix<float> &centroids);
};

// Sequential CPU Lloyd with kmeans_naive's arithmetic: fma distance, strict
// '<' argmin from FLT_MAX, Kahan-compensated FP32 sums, sum / (float)mass.
struct KmeansNaive {
    const char *name() const { return "Lloyd_Naive"; }
    void operator()(uint32_t max_iterations, const Matrix<float> &points, Matrix<float> &centroids,
                    std::vector<uint32_t> &masses, std::vector<uint32_t> &labels) const;
};

// Kahan-compensated mean of squared differences.
double mse(const std::vector<float> &a, const std::vector<float> &b);

struct ClusteringBenchmarkStats {
    explicit ClusteringBenchmarkStats(uint32_t runs = 1) : num_runs(runs) {}
    void set_dimensions(uint64_t d, uint64_t n, uint64_t k)
    {
        num_features = d;
        num_points = n;
        num_clusters = k;
    }
    void print_times() const;
    // Adds Filename/Hostname/NumFeatures/NumPoints/NumClusters parameters and
    // writes every run's three CSV files.
    std::vector<std::string> to_csv(const std::string &csv_file, const std::string &input_file);

    std::vector<uint64_t> microseconds;
    std::vector<std::shared_ptr<Measurement::Measurement>> measurements;
    uint32_t num_runs;
    uint64_t num_features = 0, num_points = 0, num_clusters = 0;
};

// Runs a strategy over host data: fresh initial centroids each run, labels
// compared against a reference clustering in verify().
class ClusteringBenchmark {
public:
    using ClClusteringFunction = std::function<std::shared_ptr<Measurement::Measurement>(
        size_t iterations, size_t features, std::shared_ptr<const std::vector<float>> points,
        std::shared_ptr<std::vector<float>> centroids, std::shared_ptr<std::vector<uint32_t>> masses,
        std::shared_ptr<std::vector<uint32_t>> labels)>;
    using InitCentroidsFunction = std::function<void(const Matrix<float> &, Matrix<float> &)>;

    ClusteringBenchmark(uint32_t runs, size_t iterations, Matrix<float> &&points);
    void initialize(size_t num_clusters, InitCentroidsFunction init);
    ClusteringBenchmarkStats run(const ClClusteringFunction &f);
    void set_verification_reference(const KmeansNaive &naive);
    uint64_t verify(const ClClusteringFunction &f);
    double mse() const;
    void print_result() const;

    const Matrix<float> &points() const { return points_; }
    const std::vector<float> &centroids() const { return centroids_.data; }
    const std::vector<uint32_t> &labels() const { return labels_; }
    const std::vector<uint32_t> &masses() const { return masses_; }

private:
    uint32_t runs_;
    size_t iterations_;
    Matrix<float> points_, centroids_, reference_centroids_;
    std::vector<uint32_t> labels_, masses_, reference_labels_, reference_masses_;
    InitCentroidsFunction init_;
};

}  // namespace Clustering
