// Named wall-clock timers and values (the interface of src/metrics/metrics.hpp:165-345),
// reported to the console by metrics_report().
#pragma once
#include <stdint.h>
#include <sys/time.h>

#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace graphchi {

struct metrics_entry {
    double t0 = 0;
};

class metrics {
  public:
    explicit metrics(std::string name = "") : name_(name) {}
    static double now() {
        timeval tv;
        gettimeofday(&tv, nullptr);
        return tv.tv_sec + tv.tv_usec * 1e-6;
    }
    metrics_entry start_time() {
        metrics_entry e;
        e.t0 = now();
        return e;
    }
    void start_time(std::string key) { open_[key] = now(); }
    void stop_time(std::string key, bool = false) {
        if (open_.count(key)) add_time(key, now() - open_[key]);
    }
    void stop_time(metrics_entry e, std::string key, bool = false) { add_time(key, now() - e.t0); }
    void add_time(std::string key, double secs) {
        times_[key].push_back(secs);
    }
    void set(std::string key, std::string value) { strings_[key] = value; }
    void set(std::string key, double value) { values_[key] = value; }
    void add(std::string key, double value) { values_[key] += value; }
    double get_time(std::string key) const {
        auto it = times_.find(key);
        double t = 0;
        if (it != times_.end())
            for (double x : it->second) t += x;
        return t;
    }
    void report() const {
        printf("=== REPORT FOR %s ===\n", name_.c_str());
        for (auto &kv : strings_) printf("%s:\t%s\n", kv.first.c_str(), kv.second.c_str());
        for (auto &kv : values_) printf("%s:\t%.6g\n", kv.first.c_str(), kv.second);
        for (auto &kv : times_) {
            double sum = 0, mn = 1e300, mx = 0;
            for (double x : kv.second) {
                sum += x;
                mn = x < mn ? x : mn;
                mx = x > mx ? x : mx;
            }
            printf("%s:\t%.6f s (count %zu, min %.6f, max %.6f, avg %.6f)\n", kv.first.c_str(), sum,
                   kv.second.size(), mn, mx, sum / kv.second.size());
        }
    }

  private:
    std::string name_;
    std::map<std::string, double> open_;
    std::map<std::string, std::vector<double>> times_;
    std::map<std::string, std::string> strings_;
    std::map<std::string, double> values_;
};

inline void metrics_report(metrics &m) { m.report(); }

}  // namespace graphchi
