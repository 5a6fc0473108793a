#include "measurement.hpp"

#include <cstdio>
#include <ctime>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>

namespace Measurement {

namespace {
const char *const kTimestampFormat = "%F-%H-%M-%S";  // measurement.cpp:25
const char *const kExperimentSuffix = "_expm";
const char *const kMeasurementsSuffix = "_mnts";
const char *const kEventsSuffix = "_evnt";

std::shared_ptr<ihipEvent_t> make_event()
{
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess)
        throw std::runtime_error("hipEventCreate failed");
    return std::shared_ptr<ihipEvent_t>(e, [](hipEvent_t x) { (void)hipEventDestroy(x); });
}

uint64_t elapsed_ns(hipEvent_t a, hipEvent_t b)
{
    (void)hipEventSynchronize(b);
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess)
        return 0;
    return ms <= 0.0f ? 0 : (uint64_t)((double)ms * 1e6);
}
}  // namespace

TimedEvent TimedEvent::begin(hipStream_t s)
{
    TimedEvent t;
    t.start = make_event();
    (void)hipEventRecord(t.start.get(), s);
    return t;
}

void TimedEvent::finish(hipStream_t s)
{
    end = make_event();
    (void)hipEventRecord(end.get(), s);
}

uint64_t TimedEvent::duration_ns() const
{
    if (!start || !end)
        return 0;
    return elapsed_ns(start.get(), end.get());
}

uint64_t DataPoint::get_value() const
{
    uint64_t value = 0;
    if (has_event_ && !events_.empty()) {
        for (const TimedEvent &e : events_)
            value += e.duration_ns();
    } else {
        for (uint64_t v : values_)
            value += v;
    }
    for (const DataPoint &c : children_)
        value += c.get_value();
    return value;
}

Measurement::Measurement() : run_date_(std::chrono::system_clock::now())
{
    set_parameter("TimeStamp", get_datetime());
}

std::string Measurement::get_unique_id() const
{
    std::random_device rand;
    std::stringstream ss;
    ss << std::dec << rand();
    return ss.str();
}

std::string Measurement::get_datetime() const
{
    char buf[32];
    std::time_t t = std::chrono::system_clock::to_time_t(run_date_);
    std::tm tmv;
    gmtime_r(&t, &tmv);
    std::strftime(buf, sizeof(buf), kTimestampFormat, &tmv);
    return buf;
}

std::string Measurement::format_filename(const std::string &base, const std::string &id,
                                         const std::string &suffix) const
{
    std::string dir, file = base;
    size_t slash = base.find_last_of('/');
    if (slash != std::string::npos) {
        dir = base.substr(0, slash + 1);
        file = base.substr(slash + 1);
    }
    std::string stem = file, ext;
    size_t dot = file.find_last_of('.');
    if (dot != std::string::npos && dot != 0) {
        stem = file.substr(0, dot);
        ext = file.substr(dot);
    }
    return dir + get_datetime() + "_" + id + "_" + stem + suffix + ext;
}

// The rows of *_mnts.csv: every top-level data point, preceded, when it
// has children (the chunk pipeline's per-chunk copies and kernels), by a
// "<name>#Sub" point holding the sum of the children's values
// (measurement/measurement.cpp:369-390).
std::deque<DataPoint> Measurement::flattened() const
{
    std::deque<DataPoint> out;
    for (const DataPoint &dp : data_points_) {
        if (!dp.children_.empty()) {
            DataPoint sub = dp.is_iterative() ? DataPoint(dp.get_iteration()) : DataPoint();
            sub.set_name(dp.get_name() + "#Sub");
            uint64_t v = 0;
            for (const DataPoint &c : dp.children_)
                v += c.get_value();
            sub.add_value() = v;
            out.push_back(sub);
        }
        out.push_back(dp);
    }
    return out;
}

// Every data point with events, children included, depth first
// (measurement/measurement.cpp:392-412).
std::deque<const DataPoint *> Measurement::with_events() const
{
    std::deque<const DataPoint *> out, stack;
    for (const DataPoint &dp : data_points_)
        stack.push_back(&dp);
    while (!stack.empty()) {
        const DataPoint *dp = stack.front();
        stack.pop_front();
        if (dp->num_events() > 0)
            out.push_back(dp);
        for (const DataPoint &c : dp->children_)
            stack.push_front(&c);
    }
    return out;
}

uint64_t Measurement::total_by_name(const std::string &name) const
{
    uint64_t v = 0;
    for (const DataPoint &dp : data_points_)
        if (dp.get_name() == name)
            v += dp.get_value();
    return v;
}

std::vector<std::string> Measurement::write_csv(const std::string &filename)
{
    const std::string id = get_unique_id();
    std::vector<std::string> files;

    std::string ef_name = format_filename(filename, id, kExperimentSuffix);
    {
        std::ofstream pf(ef_name, std::ios_base::out | std::ios::trunc);
        pf << "ExperimentID,ParameterName,Value\n";
        for (const auto &p : parameters_)
            pf << id << ',' << p.first << ',' << p.second << '\n';
    }
    files.push_back(ef_name);

    std::string mf_name = format_filename(filename, id, kMeasurementsSuffix);
    {
        std::ofstream mf(mf_name, std::ios_base::out | std::ios::trunc);
        mf << "ExperimentID,Run,TypeName,Iteration,Value\n";
        for (const DataPoint &dp : flattened()) {
            mf << id << ',' << run_ << ',' << dp.get_name() << ',';
            if (dp.is_iterative())
                mf << dp.get_iteration();
            mf << ',' << dp.get_value() << '\n';
        }
    }
    files.push_back(mf_name);

    // Event times are reported relative to the earliest recorded start event
    // (HIP exposes elapsed times between events, not absolute clocks).
    std::string ev_name = format_filename(filename, id, kEventsSuffix);
    {
        std::ofstream ef(ev_name, std::ios_base::out | std::ios::trunc);
        ef << "ExperimentID,Run,TypeName,Iteration,CommandQueueID,Queued,Submit,Start,End\n";
        const std::deque<const DataPoint *> evs = with_events();
        hipEvent_t epoch = nullptr;
        for (const DataPoint *dp : evs)
            if (dp->events_.front().start) {
                epoch = dp->events_.front().start.get();
                break;
            }
        for (const DataPoint *dp : evs) {
            for (const TimedEvent &e : dp->events_) {
                uint64_t s = (epoch && e.start) ? elapsed_ns(epoch, e.start.get()) : 0;
                uint64_t t = (epoch && e.end) ? elapsed_ns(epoch, e.end.get()) : 0;
                ef << id << ',' << run_ << ',' << dp->get_name() << ',';
                if (dp->is_iterative())
                    ef << dp->get_iteration();
                ef << ",0," << s << ',' << s << ',' << s << ',' << t << '\n';
            }
        }
    }
    files.push_back(ev_name);
    return files;
}

}  // namespace Measurement
