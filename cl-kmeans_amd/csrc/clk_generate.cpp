// Synthetic Gaussian-blob data on the host (clk_generate_blobs).
//
// Counter-based: every value is a pure function of (seed, index), so the
// output does not depend on the thread count.  Uniforms come from a 64-bit
// finaliser hash (the SplitMix64 output mix) of the counter; normals from the
// Box-Muller transform in double precision, rounded once to float.
#include "clkmeans.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform in (0, 1) from 53 hash bits.
inline double uniform(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    uint64_t h = mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + ctr);
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

inline double normal(uint64_t seed, uint64_t ctr)
{
    const double u1 = uniform(seed, 1, ctr);
    const double u2 = uniform(seed, 2, ctr);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace

namespace {

// Blob centres uniform in [lo, hi]^d; point p of blob b(p) = centre + sigma *
// N(0, 1) per feature; b(p) = p mod k (interleaved) or p / ceil(n/k)
// (grouped, the reference generator's layout, cluster_generator.cpp:178-195).
int generate(uint64_t n, uint64_t d, uint64_t k, double sigma, double lo, double hi, uint64_t seed, bool grouped,
             float *host_points, float *host_centres, int threads)
{
    if (d == 0 || k == 0 || (n > 0 && host_points == nullptr) || !(hi >= lo))
        return CLK_ERR_INVALID_ARGUMENT;
    std::vector<double> centre(d * k);
    for (uint64_t f = 0; f < d; ++f)
        for (uint64_t c = 0; c < k; ++c) {
            double u = uniform(seed, 0, f * k + c);
            centre[f * k + c] = lo + (hi - lo) * u;
        }
    if (host_centres)
        for (uint64_t i = 0; i < d * k; ++i)
            host_centres[i] = (float)centre[i];
    if (threads < 1)
        threads = 1;
    const uint64_t per_blob = (n + k - 1) / k;
    const uint64_t total = n * d;
    const uint64_t per = (total + threads - 1) / threads;
    auto work = [&](uint64_t e0, uint64_t e1) {
        for (uint64_t e = e0; e < e1; ++e) {
            const uint64_t f = e / n, p = e - f * n;
            const uint64_t c = grouped ? std::min(k - 1, p / per_blob) : p % k;
            host_points[e] = (float)(centre[f * k + c] + sigma * normal(seed, e));
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        uint64_t e0 = per * t, e1 = std::min(total, e0 + per);
        if (e0 >= e1)
            break;
        pool.emplace_back(work, e0, e1);
    }
    for (auto &th : pool)
        th.join();
    return CLK_OK;
}

}  // namespace

extern "C" int clk_generate_blobs(uint64_t n, uint64_t d, uint64_t k, double sigma, double box,
                                  uint64_t seed, float *host_points, float *host_centres,
                                  int threads)
{
    return generate(n, d, k, sigma, -box, box, seed, false, host_points, host_centres, threads);
}

extern "C" int clk_generate_points(uint64_t n, uint64_t d, uint64_t k, double radius, double domain_min,
                                   double domain_max, uint64_t seed, int grouped, float *host_points,
                                   float *host_centres, int threads)
{
    return generate(n, d, k, radius, domain_min, domain_max, seed, grouped != 0, host_points, host_centres,
                    threads);
}
