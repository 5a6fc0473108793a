// Measurement tree with the reference's CSV output (measurement/measurement.hpp:26-143,
// measurement.cpp:203-324): parameters -> *_expm.csv, data point values ->
// *_mnts.csv, kernel events -> *_evnt.csv.  Events are HIP event pairs
// (start, end) recorded around each kernel on its stream.
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace Measurement {

// A start/end event pair; durations are read (and waited for) lazily.
struct TimedEvent {
    std::shared_ptr<ihipEvent_t> start, end;
    static TimedEvent begin(hipStream_t s);
    void finish(hipStream_t s);
    uint64_t duration_ns() const;  // waits for `end`
};

class DataPoint {
    friend class Measurement;

public:
    DataPoint() = default;
    explicit DataPoint(int iteration) : iterative_(true), iteration_(iteration) {}

    DataPoint &set_name(std::string name)
    {
        name_ = std::move(name);
        return *this;
    }
    TimedEvent &add_event()
    {
        has_event_ = true;
        events_.emplace_back();
        return events_.back();
    }
    uint64_t &add_value()
    {
        values_.push_back(0);
        return values_.back();
    }
    DataPoint &create_child()
    {
        children_.emplace_back();
        return children_.back();
    }

    std::string get_name() const { return name_; }
    bool is_iterative() const { return iterative_; }
    int get_iteration() const { return iteration_; }
    // Sum of event durations (ns) if there are events, else of values, plus
    // the children's (measurement.cpp:37-60).
    uint64_t get_value() const;
    size_t num_events() const { return events_.size(); }

private:
    std::string name_;
    bool iterative_ = false;
    int iteration_ = 0;
    bool has_event_ = false;
    std::deque<TimedEvent> events_;
    std::vector<uint64_t> values_;
    std::deque<DataPoint> children_;
};

class Measurement {
public:
    Measurement();

    void set_run(int run) { run_ = run; }
    void set_parameter(const std::string &name, const std::string &value) { parameters_[name] = value; }
    const std::map<std::string, std::string> &parameters() const { return parameters_; }

    DataPoint &add_datapoint()
    {
        data_points_.emplace_back();
        return data_points_.back();
    }
    DataPoint &add_datapoint(int iteration)
    {
        data_points_.emplace_back(iteration);
        return data_points_.back();
    }

    // Writes <dir>/<UTC %F-%H-%M-%S>_<id>_<stem>_{expm,mnts,evnt}<ext>.
    // Returns the three file names.
    std::vector<std::string> write_csv(const std::string &filename);

    // Total value (ns) of all data points with the given name.
    uint64_t total_by_name(const std::string &name) const;
    const std::deque<DataPoint> &datapoints() const { return data_points_; }

private:
    std::string get_unique_id() const;
    std::string get_datetime() const;
    std::string format_filename(const std::string &base, const std::string &id,
                                const std::string &suffix) const;
    std::deque<DataPoint> flattened() const;
    std::deque<const DataPoint *> with_events() const;

    int run_ = 0;
    std::deque<DataPoint> data_points_;
    std::map<std::string, std::string> parameters_;
    std::chrono::system_clock::time_point run_date_;
};

}  // namespace Measurement
