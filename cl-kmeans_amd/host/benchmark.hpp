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
// uint64 num_points, then float32 [num_features][num_points].  A double
// matrix is filled from the float32 values (binary_format.cpp:36-43 reads a
// float per element into any matrix type).
struct BinaryFormat {
    static void read(const std::string &file, Matrix<float> &points);
    static void read(const std::string &file, Matrix<double> &points);
    static void write(const std::string &file, const Matrix<float> &points);
};

struct KmeansInitializer {
    // centroid c = point (c mod n)
    template <typename T>
    static void first_x(const Matrix<T> &points, Matrix<T> &centroids);
    // centroid c = a uniformly random point (std::random_device seeded)
    template <typename T>
    static void forgy(const Matrix<T> &points, Matrix<T> &centroids);
};

// Sequential CPU Lloyd with kmeans_naive's arithmetic in PointT: fma
// distance, strict '<' argmin from the type's max, Kahan-compensated sums,
// sum / (PointT)mass.
template <typename PointT, typename LabelT, typename MassT>
struct KmeansNaive {
    const char *name() const { return "Lloyd_Naive"; }
    void operator()(uint32_t max_iterations, const Matrix<PointT> &points, Matrix<PointT> &centroids,
                    std::vector<MassT> &masses, std::vector<LabelT> &labels) const;
};

// Kahan-compensated mean of squared differences.
template <typename T>
double mse(const std::vector<T> &a, const std::vector<T> &b);

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
// compared against a reference clustering in verify().  Instantiated for the
// float/uint32/uint32 and double/uint64/uint64 triples (bench.cpp:469-503).
template <typename PointT, typename LabelT, typename MassT>
class ClusteringBenchmark {
public:
    using ClClusteringFunction = std::function<std::shared_ptr<Measurement::Measurement>(
        size_t iterations, size_t features, std::shared_ptr<const std::vector<PointT>> points,
        std::shared_ptr<std::vector<PointT>> centroids, std::shared_ptr<std::vector<MassT>> masses,
        std::shared_ptr<std::vector<LabelT>> labels)>;
    using InitCentroidsFunction = std::function<void(const Matrix<PointT> &, Matrix<PointT> &)>;

    ClusteringBenchmark(uint32_t runs, size_t iterations, Matrix<PointT> &&points);
    void initialize(size_t num_clusters, InitCentroidsFunction init);
    ClusteringBenchmarkStats run(const ClClusteringFunction &f);
    void set_verification_reference(const KmeansNaive<PointT, LabelT, MassT> &naive);
    uint64_t verify(const ClClusteringFunction &f);
    double mse() const;
    void print_result() const;

    const Matrix<PointT> &points() const { return points_; }
    const std::vector<PointT> &centroids() const { return centroids_.data; }
    const std::vector<LabelT> &labels() const { return labels_; }
    const std::vector<MassT> &masses() const { return masses_; }

private:
    uint32_t runs_;
    size_t iterations_;
    Matrix<PointT> points_, centroids_, reference_centroids_;
    std::vector<LabelT> labels_, reference_labels_;
    std::vector<MassT> masses_, reference_masses_;
    InitCentroidsFunction init_;
};

using ClusteringBenchmarkF32 = ClusteringBenchmark<float, uint32_t, uint32_t>;
using ClusteringBenchmarkF64 = ClusteringBenchmark<double, uint64_t, uint64_t>;

}  // namespace Clustering
