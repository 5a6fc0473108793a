// clk_generate: the reference's generator CLI (generator.cpp:28-139) on the
// library's counter-based blob generator.  Usage:
//   clk_generate [--csv] [--size MiB] [--features D] [--clusters K]
//                [--radius R] [--domain_min A] [--domain_max B] [--divisor M]
//                [--seed S] [--interleaved] OUTPUT
// Writes the reference's .bin format (binary_format.cpp:20-46: u64 features,
// u64 0, u64 points, float32 points feature-major) or CSV (one point per
// line).  The point count follows the reference: size / 4 / features points,
// rounded down to a multiple of clusters and then of the divisor
// (cluster_generator.cpp:55-62).  Points are grouped by blob unless
// --interleaved; noise is centred N(0, radius).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "clkmeans.h"

namespace {

struct Options {
    bool csv = false, grouped = true;
    uint64_t megabytes = 100, features = 2, clusters = 10, divisor = 8, seed = 0;
    double radius = 10.0, domain_min = -100.0, domain_max = 100.0;
    std::string output;
};

int usage(const char *msg)
{
    if (msg)
        std::cerr << msg << "\n";
    std::cerr << "Usage: clk_generate [OPTION] [OUTPUT FILE]\nOptions:\n"
                 "  --help               Produce help message\n"
                 "  --csv                Generate CSV file (default output is binary)\n"
                 "  --size MiB           Target file size in MiB (as float-type data) (100)\n"
                 "  --features D         Number of features (aka. dimensions) (2)\n"
                 "  --clusters K         Number of clusters (10)\n"
                 "  --radius R           Cluster radius (Gaussian sigma) (10)\n"
                 "  --domain_min A       Domain space minimum for centroids (-100)\n"
                 "  --domain_max B       Domain space maximum for centroids (100)\n"
                 "  --divisor M          Number of points is a multiple of M (8)\n"
                 "  --seed S             Generator seed (0)\n"
                 "  --interleaved        Point p belongs to blob p mod K (default: grouped by blob)\n";
    return msg ? 2 : 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                std::cerr << name << " needs a value\n";
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--help")
            return usage(nullptr);
        else if (a == "--csv")
            o.csv = true;
        else if (a == "--interleaved")
            o.grouped = false;
        else if (a == "--size")
            o.megabytes = std::strtoull(next("--size"), nullptr, 10);
        else if (a == "--features")
            o.features = std::strtoull(next("--features"), nullptr, 10);
        else if (a == "--clusters")
            o.clusters = std::strtoull(next("--clusters"), nullptr, 10);
        else if (a == "--divisor")
            o.divisor = std::strtoull(next("--divisor"), nullptr, 10);
        else if (a == "--seed")
            o.seed = std::strtoull(next("--seed"), nullptr, 10);
        else if (a == "--radius")
            o.radius = std::strtod(next("--radius"), nullptr);
        else if (a == "--domain_min")
            o.domain_min = std::strtod(next("--domain_min"), nullptr);
        else if (a == "--domain_max")
            o.domain_max = std::strtod(next("--domain_max"), nullptr);
        else if (!a.empty() && a[0] == '-')
            return usage(("unknown option " + a).c_str());
        else
            o.output = a;
    }
    if (o.output.empty()) {
        std::cout << "Give me an output file!" << std::endl;
        return 1;
    }
    if (o.features == 0 || o.clusters == 0 || o.divisor == 0)
        return usage("features, clusters and divisor must be positive");
    const uint64_t floats = o.megabytes * 1024 * 1024 / sizeof(float);
    uint64_t n = floats / o.features / o.clusters * o.clusters;
    n -= n % o.divisor;
    if (n == 0)
        return usage("no points: increase --size");
    std::vector<float> pts(n * o.features);
    const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (clk_generate_points(n, o.features, o.clusters, o.radius, o.domain_min, o.domain_max, o.seed,
                            o.grouped ? 1 : 0, pts.data(), nullptr, threads) != CLK_OK) {
        std::cerr << "clk_generate: " << clk_last_error() << "\n";
        return 1;
    }
    std::ofstream fh(o.output, o.csv ? std::ios::trunc : (std::ios::binary | std::ios::trunc));
    if (!fh) {
        std::cerr << "clk_generate: cannot open " << o.output << "\n";
        return 1;
    }
    if (o.csv) {
        fh.precision(9);
        for (uint64_t p = 0; p < n; ++p) {
            for (uint64_t f = 0; f < o.features; ++f)
                fh << (f ? "," : "") << pts[f * n + p];
            fh << "\n";
        }
    } else {
        const uint64_t header[3] = {o.features, 0, n};
        fh.write(reinterpret_cast<const char *>(header), sizeof(header));
        fh.write(reinterpret_cast<const char *>(pts.data()), (std::streamsize)(pts.size() * sizeof(float)));
    }
    if (!fh) {
        std::cerr << "clk_generate: write failed\n";
        return 1;
    }
    return 0;
}
