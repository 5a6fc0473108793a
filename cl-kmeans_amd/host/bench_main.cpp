// clk_bench: the reference's bench CLI (bench.cpp:54-507) on the gfx950
// library.  Usage: clk_bench [--help] [--verbose] [--runs N] [--k K]
// [--iterations I] [--verify] [--csv FILE] --config FILE.conf INPUT.bin
// As in the reference, --runs/--k/--iterations are accepted but the .conf
// values are used (bench.cpp:201-218).  Extra: --buffer-size BYTES for the
// _buffered pipelines, --tolerance T for early convergence, --device D.
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "benchmark.hpp"
#include "configuration.hpp"
#include "kmeans_driver.hpp"

using namespace Clustering;

namespace {

struct CmdOptions {
    bool verbose = false, verify = false, csv = false;
    std::string csv_file, config_file, input_file;
    DriverOptions driver;

    int parse(int argc, char **argv)
    {
        const std::string help =
            "Usage: clk_bench [OPTION] [FILE]\nOptions:\n"
            "  --help               Produce help message\n"
            "  --verbose            Show additional information\n"
            "  --runs arg           Number of runs\n"
            "  --k arg              Number of Clusters\n"
            "  --iterations arg     Number of iterations\n"
            "  --verify             Do verification pass\n"
            "  --csv arg            Output measurements to CSV file\n"
            "  --config arg         Configuration file\n"
            "  --buffer-size arg    Bytes of points per chunk (_buffered pipelines)\n"
            "  --tolerance arg      Stop when the largest centroid move <= arg\n"
            "  --device arg         HIP device\n";
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            auto value = [&](const char *name) -> std::string {
                if (i + 1 >= argc)
                    throw std::invalid_argument(std::string("missing value for ") + name);
                return argv[++i];
            };
            if (a == "--help") {
                std::cout << help << std::endl;
                return -1;
            } else if (a == "--verbose") {
                verbose = true;
            } else if (a == "--verify") {
                verify = true;
            } else if (a == "--csv") {
                csv = true;
                csv_file = value("--csv");
            } else if (a == "--config") {
                config_file = value("--config");
            } else if (a == "--runs" || a == "--k" || a == "--iterations") {
                (void)value(a.c_str());
            } else if (a == "--buffer-size") {
                driver.buffer_size = std::stoull(value("--buffer-size"));
            } else if (a == "--tolerance") {
                driver.tolerance = std::stod(value("--tolerance"));
            } else if (a == "--device") {
                driver.device = std::stoi(value("--device"));
            } else if (!a.empty() && a[0] == '-') {
                throw std::invalid_argument("unrecognised option '" + a + "'");
            } else {
                input_file = a;
            }
        }
        if (input_file.empty()) {
            std::cerr << "No input file given" << std::endl << help;
            return -1;
        }
        if (config_file.empty()) {
            std::cerr << "No configuration file given" << std::endl << help;
            return -1;
        }
        return 0;
    }
};

// Bench<PointT, LabelT, MassT>::run (bench.cpp:200-452) for one type triple.
template <typename P, typename L, typename M>
int run_typed(const CmdOptions &opt, const ConfigurationParser &config)
{
    const BenchmarkConfiguration bm_config = config.get_benchmark_configuration();
    const KmeansConfiguration km_config = config.get_kmeans_configuration();

    Matrix<P> points;
    BinaryFormat::read(opt.input_file, points);
    const size_t n = points.rows, d = points.cols;
    ClusteringBenchmark<P, L, M> bm((uint32_t)bm_config.runs, km_config.iterations, std::move(points));
    bm.initialize(km_config.clusters, KmeansInitializer::first_x<P>);

    const bool verify = opt.verify || bm_config.verify;
    if (verify)
        bm.set_verification_reference(KmeansNaive<P, L, M>());

    auto kmeans = make_kmeans<P, L, M>(config, opt.driver);
    ClusteringBenchmarkStats bs((uint32_t)bm_config.runs);
    uint64_t wrong = 0;
    if (verify)
        wrong = bm.verify(kmeans);
    else
        bs = bm.run(kmeans);

    if (opt.verbose) {
        std::cout << "Pipeline: " << km_config.pipeline << " Types: " << km_config.point_type << " "
                  << km_config.label_type << " " << km_config.mass_type << std::endl;
        if (verify) {
            std::cout << "Centroid MSE: " << bm.mse() << std::endl;
            if (wrong == 0)
                std::cout << "Correct labels" << std::endl;
            else
                std::cout << wrong << " incorrect labels" << std::endl;
            bm.print_result();
        } else {
            bs.print_times();
        }
    }
    if (opt.csv && !verify) {
        bs.set_dimensions(d, n, km_config.clusters);
        for (const auto &f : bs.to_csv(opt.csv_file, opt.input_file))
            if (opt.verbose)
                std::cout << "wrote " << f << std::endl;
    }
    if (verify && wrong != 0)
        return 3;
    return 0;
}

// The type switch of bench.cpp:469-503.
int run(const CmdOptions &opt)
{
    ConfigurationParser config;
    config.parse_file(opt.config_file);
    if (conf_is_double(config))
        return run_typed<double, uint64_t, uint64_t>(opt, config);
    return run_typed<float, uint32_t, uint32_t>(opt, config);
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        CmdOptions opt;
        if (opt.parse(argc, argv) < 0)
            return 1;
        return run(opt);
    } catch (const std::exception &e) {
        std::cerr << "clk_bench: " << e.what() << std::endl;
        return 2;
    }
}
