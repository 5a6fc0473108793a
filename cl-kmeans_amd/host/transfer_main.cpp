// clk_transfer_bench: host<->device transfer rates by chunk size, the role of
// the reference's transfer_bench.cpp:96-206 (which streams data through its
// buffer cache and scheduler with a zero kernel per chunk).  For each buffer
// size it reports, over --size MiB of data:
//   pinned_h2d / pageable_h2d / pinned_d2h: back-to-back hipMemcpyAsync of
//     one chunk after another on one stream;
//   pipeline: the _buffered pipelines' SimpleBufferCache +
//     SingleDeviceScheduler (two device slots, copy and compute streams)
//     with a zero fill of every chunk on the compute stream.
// --verify checks every byte that moved: the device buffer after each H2D
// mode, the host buffer after the D2H, and (the pipeline then copies each
// chunk's points to its label slot instead of zeroing them, and writes the
// labels back) the host label mirror; a mismatch exits with status 3.
// Usage: clk_transfer_bench [--size MiB] [--buffer-sizes 1,4,16,64,256]
//                           [--repeat R] [--device D] [--verify]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "chunk_pipeline.hpp"
#include "device.hpp"

using namespace Clustering;

namespace {

double seconds(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Best of `repeat` runs of f (seconds).
template <typename F>
double best_of(int repeat, F f)
{
    double best = 1e30;
    for (int r = 0; r < repeat; ++r)
        best = std::min(best, f());
    return best;
}

}  // namespace

int main(int argc, char **argv)
{
    uint64_t mib = 2048;
    int repeat = 3, device = 0;
    bool verify = false;
    std::vector<uint64_t> sizes = {1, 4, 16, 64, 256};
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::cerr << a << " needs a value\n";
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--size") {
            mib = std::strtoull(next().c_str(), nullptr, 10);
        } else if (a == "--repeat") {
            repeat = std::atoi(next().c_str());
        } else if (a == "--device") {
            device = std::atoi(next().c_str());
        } else if (a == "--verify") {
            verify = true;
        } else if (a == "--buffer-sizes") {
            sizes.clear();
            std::stringstream ss(next());
            std::string tok;
            while (std::getline(ss, tok, ','))
                sizes.push_back(std::strtoull(tok.c_str(), nullptr, 10));
        } else {
            std::cerr << "Usage: clk_transfer_bench [--size MiB] [--buffer-sizes 1,4,16] [--repeat R] [--device D] "
                         "[--verify]\n";
            return a == "--help" ? 0 : 2;
        }
    }
    if (mib == 0 || sizes.empty() || repeat < 1) {
        std::cerr << "clk_transfer_bench: empty transfer\n";
        return 2;
    }
    try {
        check_hip(hipSetDevice(device), "hipSetDevice");
        const size_t bytes = mib << 20;
        const size_t n = bytes / sizeof(float);
        std::vector<float> pageable(n);
        for (size_t i = 0; i < n; ++i)
            pageable[i] = (float)(i & 1023);
        PinnedVector<float> pinned(n);
        std::memcpy(pinned.data(), pageable.data(), bytes);
        DeviceVector<float> dev(n);
        Queue q(device);
        std::vector<float> back(verify ? n : 0);
        // byte equality of `got` with the source data, or exit 3
        auto check = [&](const void *got, const char *what, uint64_t bs_mib) {
            if (std::memcmp(got, pageable.data(), bytes) != 0) {
                std::fprintf(stderr, "clk_transfer_bench: %s with %llu MiB buffers: data differ\n", what,
                             (unsigned long long)bs_mib);
                std::exit(3);
            }
        };
        auto check_device = [&](const char *what, uint64_t bs_mib) {
            check_hip(hipMemcpy(back.data(), dev.data(), bytes, hipMemcpyDeviceToHost), "verify D2H");
            check(back.data(), what, bs_mib);
            check_hip(hipMemset(dev.data(), 0, bytes), "verify clear");
        };
        std::printf("buffer_MiB,pinned_h2d_GBps,pageable_h2d_GBps,pinned_d2h_GBps,pipeline_GBps\n");
        for (uint64_t bs_mib : sizes) {
            const size_t chunk = std::max<size_t>(bs_mib << 20, sizeof(float)) / sizeof(float);
            auto copy_all = [&](float *host, hipMemcpyKind kind) {
                return best_of(repeat, [&] {
                    q.finish();
                    const auto t0 = std::chrono::steady_clock::now();
                    for (size_t off = 0; off < n; off += chunk) {
                        const size_t cn = std::min(chunk, n - off);
                        if (kind == hipMemcpyHostToDevice)
                            check_hip(hipMemcpyAsync(dev.data() + off, host + off, cn * sizeof(float), kind, q.get()),
                                      "H2D");
                        else
                            check_hip(hipMemcpyAsync(host + off, dev.data() + off, cn * sizeof(float), kind, q.get()),
                                      "D2H");
                    }
                    q.finish();
                    return seconds(t0);
                });
            };
            const double t_pin = copy_all(pinned.data(), hipMemcpyHostToDevice);
            if (verify)
                check_device("pinned H2D", bs_mib);
            const double t_page = copy_all(pageable.data(), hipMemcpyHostToDevice);
            if (verify) {
                check_device("pageable H2D", bs_mib);
                check_hip(hipMemcpy(dev.data(), pageable.data(), bytes, hipMemcpyHostToDevice), "verify H2D");
                std::memset(pinned.data(), 0, bytes);
            }
            const double t_d2h = copy_all(pinned.data(), hipMemcpyDeviceToHost);
            if (verify)
                check(pinned.data(), "pinned D2H", bs_mib);
            // the _buffered pipeline: d = 1 "points", chunks of bs bytes
            SimpleBufferCache<float, uint32_t> cache;
            cache.setup(q, pageable, 1, n, bs_mib << 20);
            const double t_pipe = best_of(repeat, [&] {
                SingleDeviceScheduler<float, uint32_t> sched(cache, q);
                q.finish();
                const auto t0 = std::chrono::steady_clock::now();
                sched.run(verify, [&](const Queue &cq, size_t cn, float *pts, uint32_t *labels) {
                    if (verify)
                        check_hip(hipMemcpyAsync(labels, pts, cn * sizeof(float), hipMemcpyDeviceToDevice, cq.get()),
                                  "chunk copy");
                    else
                        check_hip(hipMemsetAsync(pts, 0, cn * sizeof(float), cq.get()), "zero");
                });
                sched.finish();
                return seconds(t0);
            });
            if (verify) {
                std::vector<uint32_t> mirror;
                cache.copy_labels_out(mirror);
                check(mirror.data(), "pipeline", bs_mib);
            }
            const double gb = (double)bytes / 1e9;
            std::printf("%llu,%.2f,%.2f,%.2f,%.2f\n", (unsigned long long)bs_mib, gb / t_pin, gb / t_page,
                        gb / t_d2h, gb / t_pipe);
            std::fflush(stdout);
        }
    } catch (const std::exception &e) {
        std::cerr << "clk_transfer_bench: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
