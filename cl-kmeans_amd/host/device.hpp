// Device plumbing for the host layer: owning device vectors, buffer
// iterators, streams (the reference's command queues) and events.  These take
// the place of boost::compute::{vector, buffer_iterator, command_queue,
// event, wait_list} in the operator signatures, keeping their roles.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "clkmeans.h"

namespace Clustering {

// Converts a C-ABI status into the exception the reference raises in the
// same situation (std::invalid_argument for unknown strategies/arguments).
inline void check_status(int rc, const char *what)
{
    if (rc == CLK_OK)
        return;
    std::string msg = std::string(what) + ": " + clk_last_error();
    if (rc == CLK_ERR_INVALID_ARGUMENT)
        throw std::invalid_argument(msg);
    if (rc == CLK_ERR_SIZE)
        throw std::length_error(msg);
    throw std::runtime_error(msg);
}

inline void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// A HIP event, shared by copies (like boost::compute::event).
class Event {
public:
    Event() = default;
    static Event record(hipStream_t stream)
    {
        Event e;
        hipEvent_t h;
        check_hip(hipEventCreate(&h), "hipEventCreate");
        e.h_ = std::shared_ptr<ihipEvent_t>(h, [](hipEvent_t x) { (void)hipEventDestroy(x); });
        check_hip(hipEventRecord(h, stream), "hipEventRecord");
        return e;
    }
    bool valid() const { return (bool)h_; }
    hipEvent_t get() const { return h_.get(); }
    void wait() const
    {
        if (h_)
            check_hip(hipEventSynchronize(h_.get()), "hipEventSynchronize");
    }

private:
    std::shared_ptr<ihipEvent_t> h_;
};

using WaitList = std::vector<Event>;

// Make `stream` wait for every event of the list.
inline void enqueue_wait(hipStream_t stream, const WaitList &events)
{
    for (const Event &e : events)
        if (e.valid())
            check_hip(hipStreamWaitEvent(stream, e.get(), 0), "hipStreamWaitEvent");
}

// A command queue: one HIP stream on one device.
class Queue {
public:
    Queue() = default;
    explicit Queue(int device) : device_(device)
    {
        check_hip(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        check_hip(hipStreamCreate(&s), "hipStreamCreate");
        s_ = std::shared_ptr<ihipStream_t>(s, [](hipStream_t x) { (void)hipStreamDestroy(x); });
    }
    hipStream_t get() const { return s_ ? s_.get() : nullptr; }
    int device() const { return device_; }
    void finish() const { check_hip(hipStreamSynchronize(get()), "hipStreamSynchronize"); }

private:
    int device_ = 0;
    std::shared_ptr<ihipStream_t> s_;
};

// Iterator into a device buffer: base pointer + element index, as
// boost::compute::buffer_iterator (get_buffer(), get_index()).
template <typename T>
class BufferIterator {
public:
    BufferIterator() = default;
    BufferIterator(T *base, size_t index) : base_(base), index_(index) {}
    T *get_buffer() const { return base_; }
    size_t get_index() const { return index_; }
    T *get() const { return base_ + index_; }
    BufferIterator operator+(ptrdiff_t o) const { return BufferIterator(base_, index_ + o); }
    ptrdiff_t operator-(const BufferIterator &o) const { return (ptrdiff_t)index_ - (ptrdiff_t)o.index_; }

private:
    T *base_ = nullptr;
    size_t index_ = 0;
};

// Owning device vector (hipMalloc), move-only.
template <typename T>
class DeviceVector {
public:
    DeviceVector() = default;
    explicit DeviceVector(size_t n) { resize(n); }
    DeviceVector(const DeviceVector &) = delete;
    DeviceVector &operator=(const DeviceVector &) = delete;
    DeviceVector(DeviceVector &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceVector &operator=(DeviceVector &&o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_;
            o.p_ = nullptr; o.n_ = 0;
        }
        return *this;
    }
    ~DeviceVector() { release(); }

    void resize(size_t n)
    {
        if (n == n_)
            return;
        release();
        if (n)
            check_hip(hipMalloc((void **)&p_, n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    size_t size() const { return n_; }
    T *data() const { return p_; }
    BufferIterator<T> begin() const { return BufferIterator<T>(p_, 0); }
    BufferIterator<T> end() const { return BufferIterator<T>(p_, n_); }

    void copy_from_host(const T *src, size_t n, const Queue &q)
    {
        check_hip(hipMemcpyAsync(p_, src, n * sizeof(T), hipMemcpyHostToDevice, q.get()), "H2D");
    }
    void copy_to_host(T *dst, size_t n, const Queue &q) const
    {
        check_hip(hipMemcpyAsync(dst, p_, n * sizeof(T), hipMemcpyDeviceToHost, q.get()), "D2H");
    }

private:
    void release()
    {
        if (p_)
            (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *p_ = nullptr;
    size_t n_ = 0;
};

// Pinned host memory (hipHostMalloc), move-only: the staging memory of the
// out-of-core (_buffered) strategies.
template <typename T>
class PinnedVector {
public:
    PinnedVector() = default;
    explicit PinnedVector(size_t n) { resize(n); }
    PinnedVector(const PinnedVector &) = delete;
    PinnedVector &operator=(const PinnedVector &) = delete;
    ~PinnedVector() { release(); }
    void resize(size_t n)
    {
        if (n == n_)
            return;
        release();
        if (n)
            check_hip(hipHostMalloc((void **)&p_, n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
        n_ = n;
    }
    size_t size() const { return n_; }
    T *data() const { return p_; }

private:
    void release()
    {
        if (p_)
            (void)hipHostFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *p_ = nullptr;
    size_t n_ = 0;
};

// One clk_context per operator object (owns the operator's scratch, as each
// reference operator owns its tmp buffers).
class ContextHandle {
public:
    ContextHandle() = default;  // empty until assigned
    explicit ContextHandle(int device)
    {
        clk_context *c = nullptr;
        check_status(clk_context_create(device, &c), "clk_context_create");
        c_ = std::shared_ptr<clk_context>(c, [](clk_context *x) { clk_context_destroy(x); });
    }
    clk_context *get() const { return c_.get(); }

private:
    std::shared_ptr<clk_context> c_;
};

}  // namespace Clustering
