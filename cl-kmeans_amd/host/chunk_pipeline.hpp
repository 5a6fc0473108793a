// Out-of-core chunk pipeline of the _buffered strategies: the roles of the
// reference's SimpleBufferCache (simple_buffer_cache.cpp:24-705, pinned
// staging + device slots) and SingleDeviceScheduler
// (single_device_scheduler.cpp:108-259, chunk round-robin over two queues),
// rebuilt for one GPU as a double-buffered pinned hipMemcpyAsync pipeline:
//
//   host (pinned, chunk-packed)  --h2d stream-->  device slot s = i % 2
//   compute stream: wait(h2d[s], d2h[s]) -> operator(chunk i) -> comp[s]
//   device labels slot s         --d2h stream-->  host labels (pinned)
//
// Chunk i+1 uploads while chunk i computes; slot reuse is ordered by events.
// Points are repacked once, untimed, into the reference's chunk layout
// (buffer_helper.cpp:58-84: chunk b = d rows of its points, back to back), so
// every chunk is one contiguous DMA and a plain [d][chunk_n] operand.
#pragma once

#include <algorithm>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <vector>

#include "device.hpp"

namespace Clustering {

template <typename PointT, typename LabelT>
class SimpleBufferCache {
public:
    // buffer_size: bytes of points per chunk.
    void setup(const Queue &q, const std::vector<PointT> &host_points, size_t d, size_t n,
               size_t buffer_size)
    {
        if (d == 0)
            throw std::invalid_argument("SimpleBufferCache: d == 0");
        d_ = d;
        n_ = n;
        size_t cp = buffer_size / (d * sizeof(PointT));
        cp -= cp % 4;
        if (cp == 0)
            throw std::invalid_argument("SimpleBufferCache: buffer_size smaller than 4 points");
        chunk_points_ = std::min(cp, std::max<size_t>(n, 1));
        num_chunks_ = n ? (n + chunk_points_ - 1) / chunk_points_ : 0;
        points_.resize(std::max<size_t>(n * d, 1));
        labels_.resize(std::max<size_t>(n, 1));
        for (size_t b = 0; b < num_chunks_; ++b) {
            const size_t p0 = b * chunk_points_, cn = chunk_count(b);
            PointT *dst = points_.data() + p0 * d;
            for (size_t f = 0; f < d; ++f)
                std::memcpy(dst + f * cn, host_points.data() + f * n + p0, cn * sizeof(PointT));
        }
        for (int s = 0; s < 2; ++s) {
            slot_points_[s].resize(chunk_points_ * d);
            slot_labels_[s].resize(chunk_points_);
        }
        (void)q;
    }

    size_t chunk_points() const { return chunk_points_; }
    size_t num_chunks() const { return num_chunks_; }
    size_t num_features() const { return d_; }
    size_t chunk_count(size_t b) const { return std::min(chunk_points_, n_ - b * chunk_points_); }
    const PointT *host_chunk(size_t b) const { return points_.data() + b * chunk_points_ * d_; }
    LabelT *host_labels_chunk(size_t b) { return labels_.data() + b * chunk_points_; }
    PointT *slot_points(int s) const { return slot_points_[s].data(); }
    LabelT *slot_labels(int s) const { return slot_labels_[s].data(); }

    // Chunk 0 in slot 0 (blocking), for initialisers that read the first points.
    PointT *first_chunk(const Queue &q)
    {
        if (num_chunks_ == 0)
            return slot_points_[0].data();
        check_hip(hipMemcpyAsync(slot_points_[0].data(), host_chunk(0), chunk_count(0) * d_ * sizeof(PointT),
                                 hipMemcpyHostToDevice, q.get()),
                  "first chunk H2D");
        q.finish();
        return slot_points_[0].data();
    }

    void copy_labels_out(std::vector<LabelT> &out) const
    {
        out.resize(std::max(out.size(), n_));
        std::memcpy(out.data(), labels_.data(), n_ * sizeof(LabelT));
    }

private:
    size_t d_ = 0, n_ = 0, chunk_points_ = 0, num_chunks_ = 0;
    PinnedVector<PointT> points_;
    PinnedVector<LabelT> labels_;
    DeviceVector<PointT> slot_points_[2];
    DeviceVector<LabelT> slot_labels_[2];
};

template <typename PointT, typename LabelT>
class SingleDeviceScheduler {
public:
    using Runnable = std::function<void(const Queue &, size_t chunk_n, PointT *points, LabelT *labels)>;

    SingleDeviceScheduler(SimpleBufferCache<PointT, LabelT> &cache, const Queue &compute)
        : cache_(cache), compute_(compute), h2d_(compute.device()), d2h_(compute.device())
    {
        for (int s = 0; s < 2; ++s)
            for (hipEvent_t *e : {&ev_h2d_[s], &ev_comp_[s], &ev_d2h_[s]})
                check_hip(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    }
    ~SingleDeviceScheduler()
    {
        (void)hipStreamSynchronize(h2d_.get());
        (void)hipStreamSynchronize(d2h_.get());
        for (int s = 0; s < 2; ++s)
            for (hipEvent_t e : {ev_h2d_[s], ev_comp_[s], ev_d2h_[s]})
                (void)hipEventDestroy(e);
    }

    // Run `f` on every chunk in order; labels written by `f` are copied back
    // to the host mirror when write_back is set.  Returns once everything is
    // enqueued; the compute queue carries the operators' results.
    void run(bool write_back, const Runnable &f)
    {
        const size_t d = cache_.num_features();
        for (size_t i = 0; i < cache_.num_chunks(); ++i) {
            const int s = (int)(i % 2);
            const size_t cn = cache_.chunk_count(i);
            if (comp_used_[s])
                check_hip(hipStreamWaitEvent(h2d_.get(), ev_comp_[s], 0), "wait comp");
            check_hip(hipMemcpyAsync(cache_.slot_points(s), cache_.host_chunk(i), cn * d * sizeof(PointT),
                                     hipMemcpyHostToDevice, h2d_.get()),
                      "chunk H2D");
            check_hip(hipEventRecord(ev_h2d_[s], h2d_.get()), "record h2d");
            check_hip(hipStreamWaitEvent(compute_.get(), ev_h2d_[s], 0), "wait h2d");
            if (d2h_used_[s])
                check_hip(hipStreamWaitEvent(compute_.get(), ev_d2h_[s], 0), "wait d2h");
            f(compute_, cn, cache_.slot_points(s), cache_.slot_labels(s));
            check_hip(hipEventRecord(ev_comp_[s], compute_.get()), "record comp");
            comp_used_[s] = true;
            if (write_back) {
                check_hip(hipStreamWaitEvent(d2h_.get(), ev_comp_[s], 0), "wait comp");
                check_hip(hipMemcpyAsync(cache_.host_labels_chunk(i), cache_.slot_labels(s), cn * sizeof(LabelT),
                                         hipMemcpyDeviceToHost, d2h_.get()),
                          "labels D2H");
                check_hip(hipEventRecord(ev_d2h_[s], d2h_.get()), "record d2h");
                d2h_used_[s] = true;
            }
        }
    }

    void finish()
    {
        h2d_.finish();
        d2h_.finish();
        compute_.finish();
    }

private:
    SimpleBufferCache<PointT, LabelT> &cache_;
    Queue compute_, h2d_, d2h_;
    hipEvent_t ev_h2d_[2], ev_comp_[2], ev_d2h_[2];
    bool comp_used_[2] = {false, false}, d2h_used_[2] = {false, false};
};

}  // namespace Clustering
