// The Lloyd strategies: AbstractKmeans (abstract_kmeans.hpp:24-144),
// KmeansSingleStage (kmeans_single_stage.hpp:52-172), KmeansThreeStage
// (kmeans_three_stage.hpp:66-243), and the out-of-core
// KmeansSingleStageBuffered / KmeansThreeStageBuffered
// (kmeans_single_stage_buffered.hpp:49-304, kmeans_three_stage_buffered.hpp:48-415)
// over the chunk pipeline of chunk_pipeline.hpp.  One GPU: the reference's
// per-stage platform/device placement collapses to one queue.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "chunk_pipeline.hpp"
#include "operators.hpp"

namespace Clustering {

// The sum grid of a run's data set (clkmeans.h, "exact sums"), pinned on the
// context for the run and cleared after it: every op of the run sums
// exactly on the grid of all n points, so the chunked strategies give the
// unchunked sums and no op re-reads the points to find a grid.  The double
// triple sums in FP64 and has no grid.
template <typename PointT>
void pin_sum_grid(const ContextHandle &ctx, const Queue &q, size_t d, size_t n, const PointT *device_points)
{
    if constexpr (std::is_same<PointT, float>::value)
        check_status(clk_sum_grid_from_points(ctx.get(), d, n, device_points, q.get()), "clk_sum_grid_from_points");
}

// The MFMA labelings read a pinned 16-bit copy of the run's device points
// (clk_labeling_pin_points: contiguous loads in place of d strided rows)
// when there is device memory for it; without it they read the FP32 points.
// The labels are the same either way.  Returns whether a copy was pinned.
template <typename PointT>
bool pin_mfma_points(const ContextHandle &ctx, const Queue &q, const std::string &labeling, size_t d,
                     size_t n, const PointT *device_points)
{
    if constexpr (!std::is_same<PointT, float>::value) {
        return false;
    } else {
        const int id = clk_strategy_id(CLK_STAGE_LABELING, labeling.c_str());
        if ((id != CLK_LABELING_MFMA_BF16 && id != CLK_LABELING_MFMA_FP16) || n == 0)
            return false;
        return clk_labeling_pin_points(ctx.get(), id, d, n, device_points, q.get()) == CLK_OK;
    }
}

// The same from host-resident points [d][n]: per-feature max |x| over the
// finite values, on the host cores.
template <typename PointT>
void pin_sum_grid_host(const ContextHandle &ctx, size_t d, size_t n, const std::vector<PointT> &hp)
{
    if constexpr (std::is_same<PointT, float>::value) {
        std::vector<float> mx(d, 0.0f);
        const size_t nt = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), 64));
        std::vector<std::vector<float>> part(nt, std::vector<float>(d, 0.0f));
        std::vector<std::thread> th;
        const size_t per = (n + nt - 1) / nt;
        for (size_t t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const size_t p0 = std::min(n, t * per), p1 = std::min(n, p0 + per);
                for (size_t f = 0; f < d; ++f) {
                    float m = 0.0f;
                    const float *row = hp.data() + f * n;
                    for (size_t p = p0; p < p1; ++p) {
                        const float a = std::fabs(row[p]);
                        m = (a <= 3.4028234663852886e38f && a > m) ? a : m;
                    }
                    part[t][f] = m;
                }
            });
        for (auto &x : th)
            x.join();
        for (size_t t = 0; t < nt; ++t)
            for (size_t f = 0; f < d; ++f)
                mx[f] = std::max(mx[f], part[t][f]);
        check_status(clk_sum_grid_from_max(ctx.get(), d, n, mx.data()), "clk_sum_grid_from_max");
    }
}

template <typename PointT, typename LabelT, typename MassT, bool ColMajor = true>
class AbstractKmeans {
public:
    using MeasurementPtr = std::shared_ptr<Measurement::Measurement>;
    template <typename T>
    using HostVector = std::vector<T>;
    template <typename T>
    using HostVectorPtr = std::shared_ptr<std::vector<T>>;
    // Initialiser over device buffers (points [d][n], centroids [d][k]).
    using InitCentroidsFunction =
        std::function<void(const Queue &, size_t d, size_t n, size_t k, const PointT *points, PointT *centroids)>;

    AbstractKmeans() : measurement(new Measurement::Measurement) {}
    virtual ~AbstractKmeans() {}

    virtual void set_max_iterations(size_t i) { max_iterations = i; }
    virtual void set_initializer(InitCentroidsFunction f) { centroids_initializer = f; }
    virtual void set_measurement(MeasurementPtr m) { measurement = m; }
    // Stop early once the largest centroid move is <= tol (FP64, on the GPU).
    // 0 (default) runs exactly max_iterations like the reference, whose
    // kmeans.converge key is parsed but never consulted.
    virtual void set_convergence_tolerance(double tol) { convergence_tolerance = tol; }
    size_t iterations_run() const { return iterations_done; }

    virtual HostVector<PointT> const &get_centroids() const { return *host_centroids; }
    virtual HostVector<LabelT> const &get_labels() const { return *host_labels; }
    virtual HostVector<MassT> const &get_cluster_masses() const { return *host_masses; }
    virtual Measurement::Measurement const &get_measurement() const { return *measurement; }

    virtual void run() = 0;

    virtual MeasurementPtr operator()(size_t max_iter, size_t features,
                                      std::shared_ptr<const std::vector<PointT>> points,
                                      HostVectorPtr<PointT> centroids, HostVectorPtr<MassT> masses,
                                      HostVectorPtr<LabelT> labels)
    {
        max_iterations = max_iter;
        num_features = features;
        num_points = points->size() / features;
        num_clusters = centroids->size() / features;
        host_points = points;
        host_centroids = centroids;
        host_masses = masses;
        host_labels = labels;
        if (host_labels->size() < num_points)
            host_labels->resize(num_points);
        if (host_masses->size() < num_clusters)
            host_masses->resize(num_clusters);
        run();
        auto current = measurement;
        measurement.reset(new Measurement::Measurement);
        return current;
    }

protected:
    // Per-iteration convergence test (one small D2H, only when enabled).
    bool converged(const Queue &q, const DeviceVector<double> &shift)
    {
        if (convergence_tolerance <= 0.0)
            return false;
        double s = 0.0;
        check_hip(hipMemcpyAsync(&s, shift.data(), sizeof(double), hipMemcpyDeviceToHost, q.get()), "D2H shift");
        q.finish();
        return s <= convergence_tolerance;
    }

    size_t max_iterations = 0, num_features = 0, num_points = 0, num_clusters = 0;
    size_t iterations_done = 0;
    double convergence_tolerance = 0.0;
    std::shared_ptr<const std::vector<PointT>> host_points;
    HostVectorPtr<PointT> host_centroids;
    HostVectorPtr<MassT> host_masses;
    HostVectorPtr<LabelT> host_labels;
    InitCentroidsFunction centroids_initializer;
    MeasurementPtr measurement;
};

// Device-resident buffers shared by the in-core strategies.
template <typename PointT, typename LabelT, typename MassT>
struct ResidentBuffers {
    DeviceVector<PointT> points, centroids, new_centroids;
    DeviceVector<AccT> sums;
    DeviceVector<LabelT> labels;
    DeviceVector<MassT> masses;
    DeviceVector<double> shift;

    void setup(const Queue &q, size_t d, size_t n, size_t k, const std::vector<PointT> &hp,
               const std::vector<PointT> &hc, Measurement::Measurement &m)
    {
        auto &dp = m.add_datapoint().set_name("PointsH2D");
        points.resize(std::max<size_t>(n * d, 1));
        auto &te = dp.add_event();
        te = Measurement::TimedEvent::begin(q.get());
        points.copy_from_host(hp.data(), n * d, q);
        te.finish(q.get());
        centroids.resize(k * d);
        centroids.copy_from_host(hc.data(), k * d, q);
        new_centroids.resize(k * d);
        sums.resize(k * d);
        labels.resize(std::max<size_t>(n, 1));
        masses.resize(k);
        shift.resize(1);
        q.finish();
    }

    void download(const Queue &q, size_t d, size_t n, size_t k, std::vector<PointT> &hc,
                  std::vector<LabelT> &hl, std::vector<MassT> &hm)
    {
        centroids.copy_to_host(hc.data(), k * d, q);
        if (n)
            labels.copy_to_host(hl.data(), n, q);
        masses.copy_to_host(hm.data(), k, q);
        q.finish();
    }
};

template <typename PointT, typename LabelT, typename MassT, bool ColMajor = true>
class KmeansSingleStage : public AbstractKmeans<PointT, LabelT, MassT, ColMajor> {
public:
    using FusedFunction = typename FusedFactory<PointT, LabelT, MassT, ColMajor>::FusedFunction;

    void set_context(ContextHandle c) { context = c; }
    void set_queue(Queue q)
    {
        queue = q;
        this->measurement->set_parameter("FusedDevice", std::to_string(q.device()));
    }
    void set_fused(FusedConfiguration config)
    {
        FusedFactory<PointT, LabelT, MassT, ColMajor> factory;
        f_step = factory.create_step(context, config, *this->measurement);
    }

    // Each iteration is one clk_lloyd_step: zero + fused pass + FP64 divide
    // (kmeans_single_stage.hpp:95-148), then the centroid buffers swap.
    void run() override
    {
        const size_t d = this->num_features, n = this->num_points, k = this->num_clusters;
        auto &m = *this->measurement;
        buf.setup(queue, d, n, k, *this->host_points, *this->host_centroids, m);
        if (this->centroids_initializer)
            this->centroids_initializer(queue, d, n, k, buf.points.data(), buf.centroids.data());
        pin_sum_grid(context, queue, d, n, buf.points.data());
        queue.finish();

        auto t0 = std::chrono::steady_clock::now();
        WaitList none;
        this->iterations_done = 0;
        const bool track = this->convergence_tolerance > 0.0;
        for (size_t it = 0; it < this->max_iterations; ++it) {
            f_step(queue, d, n, k, buf.points.data(), buf.centroids.data(), buf.new_centroids.data(),
                   buf.sums.data(), buf.labels.data(), buf.masses.data(), track ? buf.shift.data() : nullptr,
                   m.add_datapoint((int)it), none);
            std::swap(buf.centroids, buf.new_centroids);
            ++this->iterations_done;
            if (this->converged(queue, buf.shift))
                break;
        }
        queue.finish();
        auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0);
        m.add_datapoint().set_name("TotalTime").add_value() = (uint64_t)ns.count();
        buf.download(queue, d, n, k, *this->host_centroids, *this->host_labels, *this->host_masses);
        check_status(clk_sum_grid_clear(context.get()), "clk_sum_grid_clear");
    }

private:
    ContextHandle context;
    Queue queue;
    typename FusedFactory<PointT, LabelT, MassT, ColMajor>::LloydStepFunction f_step;
    ResidentBuffers<PointT, LabelT, MassT> buf;
};

template <typename PointT, typename LabelT, typename MassT, bool ColMajor = true>
class KmeansThreeStage : public AbstractKmeans<PointT, LabelT, MassT, ColMajor> {
public:
    using LabelingFunction = typename LabelingFactory<PointT, LabelT, ColMajor>::LabelingFunction;
    using MassUpdateFunction = typename MassUpdateFactory<LabelT, MassT>::MassUpdateFunction;
    using CentroidUpdateFunction =
        typename CentroidUpdateFactory<PointT, LabelT, MassT, ColMajor>::CentroidUpdateFunction;

    void set_context(ContextHandle c) { context = c; }
    void set_queue(Queue q) { queue = q; }
    void set_labeler(LabelingConfiguration config)
    {
        LabelingFactory<PointT, LabelT, ColMajor> f;
        f_labeling = f.create(context, config, *this->measurement);
        labeling_strategy = config.strategy;
    }
    void set_mass_updater(MassUpdateConfiguration config)
    {
        MassUpdateFactory<LabelT, MassT> f;
        f_mass_update = f.create(context, config, *this->measurement);
    }
    void set_centroid_updater(CentroidUpdateConfiguration config)
    {
        CentroidUpdateFactory<PointT, LabelT, MassT, ColMajor> f;
        f_centroid_update = f.create(context, config, *this->measurement);
    }

    void run() override
    {
        const size_t d = this->num_features, n = this->num_points, k = this->num_clusters;
        auto &m = *this->measurement;
        buf.setup(queue, d, n, k, *this->host_points, *this->host_centroids, m);
        matrix_divide.prepare(context, MatrixBinaryOp<PointT, MassT>::Divide);
        if (this->centroids_initializer)
            this->centroids_initializer(queue, d, n, k, buf.points.data(), buf.centroids.data());
        pin_sum_grid(context, queue, d, n, buf.points.data());
        const bool pinned = pin_mfma_points(context, queue, labeling_strategy, d, n, buf.points.data());
        queue.finish();

        auto t0 = std::chrono::steady_clock::now();
        WaitList none;
        this->iterations_done = 0;
        for (size_t it = 0; it < this->max_iterations; ++it) {
            auto pb = buf.points.begin(), pe = buf.points.begin() + n * d;
            auto lb = buf.labels.begin(), le = buf.labels.begin() + n;
            f_labeling(queue, d, n, k, pb, pe, buf.centroids.begin(), buf.centroids.end(), lb, le,
                       m.add_datapoint((int)it), none);
            check_status(clk_fill_zero(context.get(), buf.masses.data(), k * sizeof(MassT), queue.get()), "fill");
            check_status(clk_fill_zero(context.get(), buf.sums.data(), k * d * sizeof(AccT), queue.get()), "fill");
            f_mass_update(queue, n, k, lb, le, buf.masses.begin(), buf.masses.end(), m.add_datapoint((int)it),
                          none);
            f_centroid_update(queue, d, n, k, pb, pe, buf.sums.begin(), buf.sums.end(), lb, le,
                              buf.masses.begin(), buf.masses.end(), m.add_datapoint((int)it), none);
            const bool track = this->convergence_tolerance > 0.0;
            matrix_divide.row(queue, d, k, buf.sums.begin(), buf.sums.end(), buf.masses.begin(),
                              buf.masses.end(), buf.new_centroids.begin(), buf.new_centroids.end(),
                              m.add_datapoint((int)it), none, track ? buf.centroids.data() : nullptr,
                              track ? buf.shift.data() : nullptr);
            std::swap(buf.centroids, buf.new_centroids);
            ++this->iterations_done;
            if (this->converged(queue, buf.shift))
                break;
        }
        queue.finish();
        auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0);
        m.add_datapoint().set_name("TotalTime").add_value() = (uint64_t)ns.count();
        buf.download(queue, d, n, k, *this->host_centroids, *this->host_labels, *this->host_masses);
        if (pinned)
            check_status(clk_labeling_unpin_points(context.get()), "clk_labeling_unpin_points");
        check_status(clk_sum_grid_clear(context.get()), "clk_sum_grid_clear");
    }

private:
    std::string labeling_strategy = "unroll_vector";
    ContextHandle context;
    Queue queue;
    LabelingFunction f_labeling;
    MassUpdateFunction f_mass_update;
    CentroidUpdateFunction f_centroid_update;
    MatrixBinaryOp<PointT, MassT> matrix_divide;
    ResidentBuffers<PointT, LabelT, MassT> buf;
};

// Buffers of the out-of-core strategies: only centroids, sums and masses are
// device-resident; points and labels stream through the chunk pipeline.
template <typename PointT, typename LabelT, typename MassT>
struct StreamingState {
    DeviceVector<PointT> centroids, new_centroids;
    DeviceVector<AccT> sums;
    DeviceVector<MassT> masses;
    DeviceVector<double> shift;
};

template <typename PointT, typename LabelT, typename MassT, bool ColMajor = true>
class KmeansSingleStageBuffered : public AbstractKmeans<PointT, LabelT, MassT, ColMajor> {
public:
    using FusedFunction = typename FusedFactory<PointT, LabelT, MassT, ColMajor>::FusedFunction;

    void set_context(ContextHandle c) { context = c; }
    void set_queue(Queue q) { queue = q; }
    void set_fused(FusedConfiguration config)
    {
        FusedFactory<PointT, LabelT, MassT, ColMajor> factory;
        f_fused = factory.create(context, config, *this->measurement);
    }
    // Bytes of points per chunk (the reference's buffer_size, 16 MiB there,
    // kmeans_single_stage_buffered.hpp:333).
    void set_buffer_size(size_t bytes) { buffer_size = bytes; }

    void run() override
    {
        const size_t d = this->num_features, n = this->num_points, k = this->num_clusters;
        auto &m = *this->measurement;
        SimpleBufferCache<PointT, LabelT> cache;
        cache.setup(queue, *this->host_points, d, n, buffer_size);
        SingleDeviceScheduler<PointT, LabelT> scheduler(cache, queue);
        st.centroids.resize(k * d);
        st.centroids.copy_from_host(this->host_centroids->data(), k * d, queue);
        st.new_centroids.resize(k * d);
        st.sums.resize(k * d);
        st.masses.resize(k);
        st.shift.resize(1);
        matrix_divide.prepare(context, MatrixBinaryOp<PointT, MassT>::Divide);
        if (this->centroids_initializer)
            this->centroids_initializer(queue, d, std::min(n, cache.chunk_points()), k, cache.first_chunk(queue),
                                        st.centroids.data());
        pin_sum_grid_host(context, d, n, *this->host_points);
        queue.finish();
        this->measurement->set_parameter("BufferSize", std::to_string(buffer_size));

        auto t0 = std::chrono::steady_clock::now();
        WaitList none;
        this->iterations_done = 0;
        for (size_t it = 0; it < this->max_iterations; ++it) {
            check_status(clk_fill_zero(context.get(), st.masses.data(), k * sizeof(MassT), queue.get()), "fill");
            check_status(clk_fill_zero(context.get(), st.sums.data(), k * d * sizeof(AccT), queue.get()), "fill");
            auto &dp = m.add_datapoint((int)it);
            scheduler.run(true, [&](const Queue &q, size_t cn, PointT *pts, LabelT *lab) {
                f_fused(q, d, cn, k, BufferIterator<PointT>(pts, 0), BufferIterator<PointT>(pts, cn * d),
                        st.centroids.begin(), st.centroids.end(), st.sums.begin(), st.sums.end(),
                        BufferIterator<LabelT>(lab, 0), BufferIterator<LabelT>(lab, cn), st.masses.begin(),
                        st.masses.end(), dp.create_child(), none);
            });
            const bool track = this->convergence_tolerance > 0.0;
            matrix_divide.row(queue, d, k, st.sums.begin(), st.sums.end(), st.masses.begin(),
                              st.masses.end(), st.new_centroids.begin(), st.new_centroids.end(),
                              m.add_datapoint((int)it), none, track ? st.centroids.data() : nullptr,
                              track ? st.shift.data() : nullptr);
            std::swap(st.centroids, st.new_centroids);
            ++this->iterations_done;
            if (this->converged(queue, st.shift))
                break;
        }
        queue.finish();
        auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0);
        m.add_datapoint().set_name("TotalTime").add_value() = (uint64_t)ns.count();
        st.centroids.copy_to_host(this->host_centroids->data(), k * d, queue);
        st.masses.copy_to_host(this->host_masses->data(), k, queue);
        scheduler.finish();
        cache.copy_labels_out(*this->host_labels);
        check_status(clk_sum_grid_clear(context.get()), "clk_sum_grid_clear");
    }

private:
    ContextHandle context;
    Queue queue;
    FusedFunction f_fused;
    MatrixBinaryOp<PointT, MassT> matrix_divide;
    StreamingState<PointT, LabelT, MassT> st;
    size_t buffer_size = 256ull << 20;
};

template <typename PointT, typename LabelT, typename MassT, bool ColMajor = true>
class KmeansThreeStageBuffered : public AbstractKmeans<PointT, LabelT, MassT, ColMajor> {
public:
    using LabelingFunction = typename LabelingFactory<PointT, LabelT, ColMajor>::LabelingFunction;
    using MassUpdateFunction = typename MassUpdateFactory<LabelT, MassT>::MassUpdateFunction;
    using CentroidUpdateFunction =
        typename CentroidUpdateFactory<PointT, LabelT, MassT, ColMajor>::CentroidUpdateFunction;

    void set_context(ContextHandle c) { context = c; }
    void set_queue(Queue q) { queue = q; }
    void set_labeler(LabelingConfiguration config)
    {
        LabelingFactory<PointT, LabelT, ColMajor> f;
        f_labeling = f.create(context, config, *this->measurement);
    }
    void set_mass_updater(MassUpdateConfiguration config)
    {
        MassUpdateFactory<LabelT, MassT> f;
        f_mass_update = f.create(context, config, *this->measurement);
    }
    void set_centroid_updater(CentroidUpdateConfiguration config)
    {
        CentroidUpdateFactory<PointT, LabelT, MassT, ColMajor> f;
        f_centroid_update = f.create(context, config, *this->measurement);
    }
    void set_buffer_size(size_t bytes) { buffer_size = bytes; }

    // One chunk pass per iteration: each resident chunk is labelled, then its
    // masses and feature sums are added, so points cross the host link once
    // per iteration (the reference streams them twice, once for labeling and
    // once for the centroid update, kmeans_three_stage_buffered.hpp:211-339).
    void run() override
    {
        const size_t d = this->num_features, n = this->num_points, k = this->num_clusters;
        auto &m = *this->measurement;
        SimpleBufferCache<PointT, LabelT> cache;
        cache.setup(queue, *this->host_points, d, n, buffer_size);
        SingleDeviceScheduler<PointT, LabelT> scheduler(cache, queue);
        st.centroids.resize(k * d);
        st.centroids.copy_from_host(this->host_centroids->data(), k * d, queue);
        st.new_centroids.resize(k * d);
        st.sums.resize(k * d);
        st.masses.resize(k);
        st.shift.resize(1);
        matrix_divide.prepare(context, MatrixBinaryOp<PointT, MassT>::Divide);
        if (this->centroids_initializer)
            this->centroids_initializer(queue, d, std::min(n, cache.chunk_points()), k, cache.first_chunk(queue),
                                        st.centroids.data());
        pin_sum_grid_host(context, d, n, *this->host_points);
        queue.finish();
        this->measurement->set_parameter("BufferSize", std::to_string(buffer_size));

        auto t0 = std::chrono::steady_clock::now();
        WaitList none;
        this->iterations_done = 0;
        for (size_t it = 0; it < this->max_iterations; ++it) {
            check_status(clk_fill_zero(context.get(), st.masses.data(), k * sizeof(MassT), queue.get()), "fill");
            check_status(clk_fill_zero(context.get(), st.sums.data(), k * d * sizeof(AccT), queue.get()), "fill");
            auto &dp = m.add_datapoint((int)it);
            scheduler.run(true, [&](const Queue &q, size_t cn, PointT *pts, LabelT *lab) {
                BufferIterator<PointT> pb(pts, 0), pe(pts, cn * d);
                BufferIterator<LabelT> lb(lab, 0), le(lab, cn);
                f_labeling(q, d, cn, k, pb, pe, st.centroids.begin(), st.centroids.end(), lb, le,
                           dp.create_child(), none);
                f_mass_update(q, cn, k, lb, le, st.masses.begin(), st.masses.end(), dp.create_child(), none);
                f_centroid_update(q, d, cn, k, pb, pe, st.sums.begin(), st.sums.end(), lb, le,
                                  st.masses.begin(), st.masses.end(), dp.create_child(), none);
            });
            const bool track = this->convergence_tolerance > 0.0;
            matrix_divide.row(queue, d, k, st.sums.begin(), st.sums.end(), st.masses.begin(),
                              st.masses.end(), st.new_centroids.begin(), st.new_centroids.end(),
                              m.add_datapoint((int)it), none, track ? st.centroids.data() : nullptr,
                              track ? st.shift.data() : nullptr);
            std::swap(st.centroids, st.new_centroids);
            ++this->iterations_done;
            if (this->converged(queue, st.shift))
                break;
        }
        queue.finish();
        auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0);
        m.add_datapoint().set_name("TotalTime").add_value() = (uint64_t)ns.count();
        st.centroids.copy_to_host(this->host_centroids->data(), k * d, queue);
        st.masses.copy_to_host(this->host_masses->data(), k, queue);
        scheduler.finish();
        cache.copy_labels_out(*this->host_labels);
        check_status(clk_sum_grid_clear(context.get()), "clk_sum_grid_clear");
    }

private:
    ContextHandle context;
    Queue queue;
    LabelingFunction f_labeling;
    MassUpdateFunction f_mass_update;
    CentroidUpdateFunction f_centroid_update;
    MatrixBinaryOp<PointT, MassT> matrix_divide;
    StreamingState<PointT, LabelT, MassT> st;
    size_t buffer_size = 256ull << 20;
};

}  // namespace Clustering
