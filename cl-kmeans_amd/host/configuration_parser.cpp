#include "configuration.hpp"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <set>
#include <sstream>
#include <stdexcept>

namespace Clustering {

namespace {

// The registered keys (configuration_parser.cpp:43-111).
const std::set<std::string> &known_keys()
{
    static const std::set<std::string> k = {
        "benchmark.runs", "benchmark.verify",
        "kmeans.clusters", "kmeans.pipeline", "kmeans.iterations", "kmeans.converge",
        "kmeans.types.point", "kmeans.types.label", "kmeans.types.mass",
        "kmeans.labeling.platform", "kmeans.labeling.device", "kmeans.labeling.strategy",
        "kmeans.labeling.global_size", "kmeans.labeling.local_size", "kmeans.labeling.vector_length",
        "kmeans.labeling.unroll_clusters_length", "kmeans.labeling.unroll_features_length",
        "kmeans.mass_update.platform", "kmeans.mass_update.device", "kmeans.mass_update.strategy",
        "kmeans.mass_update.global_size", "kmeans.mass_update.local_size",
        "kmeans.mass_update.vector_length",
        "kmeans.centroid_update.platform", "kmeans.centroid_update.device",
        "kmeans.centroid_update.strategy", "kmeans.centroid_update.global_size",
        "kmeans.centroid_update.local_size", "kmeans.centroid_update.local_features",
        "kmeans.centroid_update.thread_features", "kmeans.centroid_update.vector_length",
        "kmeans.fused.platform", "kmeans.fused.device", "kmeans.fused.strategy",
        "kmeans.fused.global_size", "kmeans.fused.local_size", "kmeans.fused.vector_length",
    };
    return k;
}

bool is_multi(const std::string &key)
{
    return key.size() > 12 && (key.compare(key.size() - 12, 12, ".global_size") == 0 ||
                               key.compare(key.size() - 11, 11, ".local_size") == 0);
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a]))
        ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1]))
        --b;
    return s.substr(a, b - a);
}

size_t to_size(const std::string &key, const std::string &v)
{
    if (v.empty() || !std::all_of(v.begin(), v.end(), [](char c) { return std::isdigit((unsigned char)c); }))
        throw std::invalid_argument("the argument ('" + v + "') for option '" + key + "' is invalid");
    return std::stoull(v);
}

bool to_bool(const std::string &key, const std::string &v)
{
    // boost::program_options accepts true/false, yes/no, on/off, 1/0.
    std::string l = v;
    std::transform(l.begin(), l.end(), l.begin(), [](unsigned char c) { return std::tolower(c); });
    if (l == "true" || l == "yes" || l == "on" || l == "1")
        return true;
    if (l == "false" || l == "no" || l == "off" || l == "0")
        return false;
    throw std::invalid_argument("the argument ('" + v + "') for option '" + key + "' is invalid");
}

void fill_sizes(const std::string &key, const std::vector<std::string> &vals, size_t out[3])
{
    std::vector<size_t> v;
    for (const std::string &s : vals) {
        std::stringstream ss(s);
        std::string tok;
        while (ss >> tok)
            v.push_back(to_size(key, tok));
    }
    if (v.empty() || v.size() > 3)
        throw std::invalid_argument("option '" + key + "' takes 1 to 3 sizes");
    out[0] = v[0];
    out[1] = v.size() > 1 ? v[1] : 1;
    out[2] = v.size() == 3 ? v[2] : 1;
}

}  // namespace

void ConfigurationParser::parse_file(const std::string &file)
{
    std::ifstream fh(file);
    if (!fh.is_open())
        throw std::invalid_argument("Could not open file \"" + file + "\"");
    std::stringstream ss;
    ss << fh.rdbuf();
    parse_string(ss.str());
}

void ConfigurationParser::parse_string(const std::string &text)
{
    std::stringstream in(text);
    std::string line, section;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        size_t hash = line.find('#');
        if (hash != std::string::npos)
            line = line.substr(0, hash);
        line = trim(line);
        if (line.empty())
            continue;
        if (line.front() == '[') {
            if (line.back() != ']')
                throw std::invalid_argument("malformed section header at line " + std::to_string(lineno));
            section = trim(line.substr(1, line.size() - 2));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos)
            throw std::invalid_argument("expected 'key = value' at line " + std::to_string(lineno));
        std::string key = trim(line.substr(0, eq));
        std::string val = trim(line.substr(eq + 1));
        std::string full = section.empty() ? key : section + "." + key;
        if (!known_keys().count(full))
            throw std::invalid_argument("unrecognised option '" + full + "'");
        auto &slot = values_[full];
        if (!slot.empty() && !is_multi(full))
            throw std::invalid_argument("option '" + full + "' cannot be specified more than once");
        slot.push_back(val);
    }
}

const std::vector<std::string> *ConfigurationParser::find(const std::string &key) const
{
    auto it = values_.find(key);
    return it == values_.end() ? nullptr : &it->second;
}

BenchmarkConfiguration ConfigurationParser::get_benchmark_configuration() const
{
    BenchmarkConfiguration c;
    if (auto v = find("benchmark.runs")) c.runs = to_size("benchmark.runs", v->front());
    if (auto v = find("benchmark.verify")) c.verify = to_bool("benchmark.verify", v->front());
    return c;
}

KmeansConfiguration ConfigurationParser::get_kmeans_configuration() const
{
    KmeansConfiguration c;
    if (auto v = find("kmeans.clusters")) c.clusters = to_size("kmeans.clusters", v->front());
    if (auto v = find("kmeans.pipeline")) c.pipeline = v->front();
    if (auto v = find("kmeans.iterations")) c.iterations = to_size("kmeans.iterations", v->front());
    if (auto v = find("kmeans.converge")) c.converge = to_bool("kmeans.converge", v->front());
    if (auto v = find("kmeans.types.point")) c.point_type = v->front();
    if (auto v = find("kmeans.types.label")) c.label_type = v->front();
    if (auto v = find("kmeans.types.mass")) c.mass_type = v->front();
    return c;
}

#define CLK_STAGE_COMMON(prefix)                                                       \
    if (auto v = find(prefix ".platform")) c.platform = to_size(prefix ".platform", v->front()); \
    if (auto v = find(prefix ".device")) c.device = to_size(prefix ".device", v->front());       \
    if (auto v = find(prefix ".strategy")) c.strategy = v->front();                              \
    if (auto v = find(prefix ".global_size")) fill_sizes(prefix ".global_size", *v, c.global_size); \
    if (auto v = find(prefix ".local_size")) fill_sizes(prefix ".local_size", *v, c.local_size);    \
    if (auto v = find(prefix ".vector_length"))                                                  \
        c.vector_length = to_size(prefix ".vector_length", v->front());

LabelingConfiguration ConfigurationParser::get_labeling_configuration() const
{
    LabelingConfiguration c;
    CLK_STAGE_COMMON("kmeans.labeling")
    if (auto v = find("kmeans.labeling.unroll_clusters_length"))
        c.unroll_clusters_length = to_size("kmeans.labeling.unroll_clusters_length", v->front());
    if (auto v = find("kmeans.labeling.unroll_features_length"))
        c.unroll_features_length = to_size("kmeans.labeling.unroll_features_length", v->front());
    return c;
}

MassUpdateConfiguration ConfigurationParser::get_mass_update_configuration() const
{
    MassUpdateConfiguration c;
    CLK_STAGE_COMMON("kmeans.mass_update")
    return c;
}

CentroidUpdateConfiguration ConfigurationParser::get_centroid_update_configuration() const
{
    CentroidUpdateConfiguration c;
    CLK_STAGE_COMMON("kmeans.centroid_update")
    if (auto v = find("kmeans.centroid_update.local_features"))
        c.local_features = to_size("kmeans.centroid_update.local_features", v->front());
    if (auto v = find("kmeans.centroid_update.thread_features"))
        c.thread_features = to_size("kmeans.centroid_update.thread_features", v->front());
    return c;
}

FusedConfiguration ConfigurationParser::get_fused_configuration() const
{
    FusedConfiguration c;
    CLK_STAGE_COMMON("kmeans.fused")
    return c;
}

#undef CLK_STAGE_COMMON

}  // namespace Clustering
