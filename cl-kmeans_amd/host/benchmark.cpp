#include "benchmark.hpp"

#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <random>
#include <stdexcept>

namespace Clustering {

void BinaryFormat::read(const std::string &file, Matrix<float> &points)
{
    std::ifstream fh(file, std::ios::binary);
    if (!fh.is_open())
        throw std::invalid_argument("Could not open file \"" + file + "\"");
    uint64_t hdr[3];
    fh.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    if (!fh)
        throw std::invalid_argument("\"" + file + "\": truncated header");
    const uint64_t d = hdr[0], k = hdr[1], n = hdr[2];
    if (k != 0)  // ground-truth centroids are not supported (binary_format.cpp:33-34)
        throw std::invalid_argument("\"" + file + "\": num_clusters field must be 0");
    points.resize(n, d);
    fh.read(reinterpret_cast<char *>(points.data.data()), (std::streamsize)(n * d * sizeof(float)));
    if (!fh)
        throw std::invalid_argument("\"" + file + "\": truncated point data");
}

void BinaryFormat::read(const std::string &file, Matrix<double> &points)
{
    Matrix<float> f;
    read(file, f);
    points.resize(f.rows, f.cols);
    for (size_t i = 0; i < f.data.size(); ++i)
        points.data[i] = f.data[i];
}

void BinaryFormat::write(const std::string &file, const Matrix<float> &points)
{
    std::ofstream fh(file, std::ios::binary | std::ios::trunc);
    uint64_t hdr[3] = {points.cols, 0, points.rows};
    fh.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    fh.write(reinterpret_cast<const char *>(points.data.data()),
             (std::streamsize)(points.data.size() * sizeof(float)));
}

template <typename T>
void KmeansInitializer::first_x(const Matrix<T> &points, Matrix<T> &centroids)
{
    for (size_t f = 0; f < centroids.cols; ++f)
        for (size_t c = 0; c < centroids.rows; ++c)
            centroids(c, f) = points(c % points.rows, f);
}

template <typename T>
void KmeansInitializer::forgy(const Matrix<T> &points, Matrix<T> &centroids)
{
    std::random_device rand;
    for (size_t c = 0; c < centroids.rows; ++c) {
        size_t p = rand() % points.rows;
        for (size_t f = 0; f < centroids.cols; ++f)
            centroids(c, f) = points(p, f);
    }
}

template <typename PointT, typename LabelT, typename MassT>
void KmeansNaive<PointT, LabelT, MassT>::operator()(uint32_t max_iterations, const Matrix<PointT> &points,
                                                    Matrix<PointT> &centroids, std::vector<MassT> &masses,
                                                    std::vector<LabelT> &labels) const
{
    const size_t n = points.rows, d = points.cols, k = centroids.rows;
    labels.assign(n, 0);
    masses.assign(k, 0);
    std::vector<PointT> comp(k * d);
    for (uint32_t it = 0; it < max_iterations; ++it) {
        for (size_t p = 0; p < n; ++p) {
            PointT best = std::numeric_limits<PointT>::max();
            LabelT lab = 0;
            for (size_t c = 0; c < k; ++c) {
                PointT dist = 0;
                for (size_t f = 0; f < d; ++f) {
                    PointT t = points(p, f) - centroids(c, f);
                    dist = std::fma(t, t, dist);
                }
                if (dist < best) {
                    best = dist;
                    lab = (LabelT)c;
                }
            }
            labels[p] = lab;
        }
        std::fill(masses.begin(), masses.end(), MassT(0));
        std::fill(centroids.data.begin(), centroids.data.end(), PointT(0));
        std::fill(comp.begin(), comp.end(), PointT(0));
        for (size_t p = 0; p < n; ++p) {
            const LabelT c = labels[p];
            masses[c] += 1;
            for (size_t f = 0; f < d; ++f) {
                PointT &s = centroids(c, f);
                PointT &cm = comp[f * k + c];
                PointT y = points(p, f) - cm;
                PointT t = s + y;
                cm = (t - s) - y;
                s = t;
            }
        }
        for (size_t f = 0; f < d; ++f)
            for (size_t c = 0; c < k; ++c)
                centroids(c, f) = centroids(c, f) / (PointT)masses[c];
    }
}

template <typename T>
double mse(const std::vector<T> &a, const std::vector<T> &b)
{
    double sum = 0.0, comp = 0.0;
    uint64_t count = 0;
    for (size_t i = 0; i < a.size() && i < b.size(); ++i) {
        double x = (double)a[i] - (double)b[i];
        double y = x * x - comp;
        double t = sum + y;
        comp = (t - sum) - y;
        sum = t;
        ++count;
    }
    return count ? sum / (double)count : 0.0;
}

void ClusteringBenchmarkStats::print_times() const
{
    std::cout << num_runs << " runs, in µs: [";
    for (size_t r = 0; r < microseconds.size(); ++r) {
        std::cout << microseconds[r];
        if (r + 1 != microseconds.size())
            std::cout << ", ";
    }
    std::cout << "]" << std::endl;
}

std::vector<std::string> ClusteringBenchmarkStats::to_csv(const std::string &csv_file,
                                                          const std::string &input_file)
{
    char host[256] = {0};
    gethostname(host, sizeof(host) - 1);
    std::string base = input_file.substr(input_file.find_last_of('/') == std::string::npos
                                             ? 0
                                             : input_file.find_last_of('/') + 1);
    std::vector<std::string> files;
    for (auto &m : measurements) {
        m->set_parameter("Version", "cl-kmeans_amd");
        m->set_parameter("Filename", base);
        m->set_parameter("Hostname", host);
        m->set_parameter("NumFeatures", std::to_string(num_features));
        m->set_parameter("NumPoints", std::to_string(num_points));
        m->set_parameter("NumClusters", std::to_string(num_clusters));
        for (auto &f : m->write_csv(csv_file))
            files.push_back(f);
    }
    return files;
}

template <typename PointT, typename LabelT, typename MassT>
ClusteringBenchmark<PointT, LabelT, MassT>::ClusteringBenchmark(uint32_t runs, size_t iterations,
                                                              Matrix<PointT> &&points)
    : runs_(runs), iterations_(iterations), points_(std::move(points))
{
}

template <typename PointT, typename LabelT, typename MassT>
void ClusteringBenchmark<PointT, LabelT, MassT>::initialize(size_t num_clusters, InitCentroidsFunction init)
{
    init_ = init;
    centroids_.resize(num_clusters, points_.cols);
    masses_.assign(num_clusters, 0);
    labels_.assign(points_.rows, 0);
}

template <typename PointT, typename LabelT, typename MassT>
ClusteringBenchmarkStats ClusteringBenchmark<PointT, LabelT, MassT>::run(const ClClusteringFunction &f)
{
    ClusteringBenchmarkStats bs(runs_);
    bs.set_dimensions(points_.cols, points_.rows, centroids_.rows);
    // Non-owning views of the benchmark's vectors (clustering_benchmark.cpp:178-193).
    std::shared_ptr<const std::vector<PointT>> pts(&points_.data, [](const std::vector<PointT> *) {});
    std::shared_ptr<std::vector<PointT>> cen(&centroids_.data, [](std::vector<PointT> *) {});
    std::shared_ptr<std::vector<MassT>> ms(&masses_, [](std::vector<MassT> *) {});
    std::shared_ptr<std::vector<LabelT>> lb(&labels_, [](std::vector<LabelT> *) {});
    for (uint32_t r = 0; r < runs_; ++r) {
        init_(points_, centroids_);
        auto t0 = std::chrono::steady_clock::now();
        auto m = f(iterations_, points_.cols, pts, cen, ms, lb);
        auto us = std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0);
        bs.microseconds.push_back((uint64_t)us.count());
        m->set_run((int)r);
        bs.measurements.push_back(m);
    }
    return bs;
}

template <typename PointT, typename LabelT, typename MassT>
void ClusteringBenchmark<PointT, LabelT, MassT>::set_verification_reference(
    const KmeansNaive<PointT, LabelT, MassT> &naive)
{
    reference_centroids_.resize(centroids_.rows, centroids_.cols);
    init_(points_, reference_centroids_);
    naive((uint32_t)iterations_, points_, reference_centroids_, reference_masses_, reference_labels_);
}

template <typename PointT, typename LabelT, typename MassT>
uint64_t ClusteringBenchmark<PointT, LabelT, MassT>::verify(const ClClusteringFunction &f)
{
    run(f);
    uint64_t wrong = 0;
    for (size_t i = 0; i < labels_.size() && i < reference_labels_.size(); ++i)
        wrong += labels_[i] != reference_labels_[i];
    return wrong;
}

template <typename PointT, typename LabelT, typename MassT>
double ClusteringBenchmark<PointT, LabelT, MassT>::mse() const
{
    return Clustering::mse(centroids_.data, reference_centroids_.data);
}

template <typename PointT, typename LabelT, typename MassT>
void ClusteringBenchmark<PointT, LabelT, MassT>::print_result() const
{
    auto dump = [](const char *title, const Matrix<PointT> &m) {
        std::cout << title << std::endl;
        for (size_t r = 0; r < m.rows; ++r) {
            for (size_t c = 0; c < m.cols; ++c)
                std::cout << m(r, c) << " ";
            std::cout << std::endl;
        }
    };
    dump("Centroids", centroids_);
    dump("Centroids (Reference)", reference_centroids_);
    std::cout << "Cluster Mass" << std::endl;
    for (auto m : masses_)
        std::cout << m << " ";
    std::cout << std::endl << "Cluster Mass (Reference)" << std::endl;
    for (auto m : reference_masses_)
        std::cout << m << " ";
    std::cout << std::endl;
}

template void KmeansInitializer::first_x(const Matrix<float> &, Matrix<float> &);
template void KmeansInitializer::first_x(const Matrix<double> &, Matrix<double> &);
template void KmeansInitializer::forgy(const Matrix<float> &, Matrix<float> &);
template void KmeansInitializer::forgy(const Matrix<double> &, Matrix<double> &);
template struct KmeansNaive<float, uint32_t, uint32_t>;
template struct KmeansNaive<double, uint64_t, uint64_t>;
template double mse(const std::vector<float> &, const std::vector<float> &);
template double mse(const std::vector<double> &, const std::vector<double> &);
template class ClusteringBenchmark<float, uint32_t, uint32_t>;
template class ClusteringBenchmark<double, uint64_t, uint64_t>;

}  // namespace Clustering
