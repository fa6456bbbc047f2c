// ScopedTimer.hpp -- timer for a named scope with aggregated statistics, the
// API and output of util/ScopedTimer.hpp:34-149: when ScopedTimer::enabled is
// true each timer prints its run time and running totals when it goes out of
// scope, and printOverview() lists every name by total time.
//
// The time is measured on the device: a pair of HIP events recorded on the
// thread's timing stream when the scope opens and closes (tsp_timer_*), so a
// scope that enqueues device steps reports the time the stream spent from the
// scope's start to its end, not the host's launch time and not work queued
// before the scope. Host work inside a scope counts too (the stream's clock
// runs while it is idle). The timing stream is the stream of the most recently
// constructed integrator of the thread, or the one a StreamGuard names; with
// none, or where no device is present, the scope reads the host clock.
#pragma once

#include "treesap_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace treesap {

class ScopedTimer {
    using Clock = std::chrono::steady_clock;

    struct Stats {
        double total_us = 0.0;
        double sum_sq_us = 0.0;
        std::size_t count = 0;
        double mean_us() const { return total_us / count; }
        double stdev_us() const {
            const double m = mean_us();
            const double var = sum_sq_us / count - m * m;
            return var > 0.0 ? std::sqrt(var) : 0.0;
        }
    };

  public:
    explicit ScopedTimer(std::string name) : name_(std::move(name)) {
        if (!enabled) return;
        active_ = true;
        if (void *s = timing_stream()) {
            dev_stream_ = s;
            dev_ = acquire(s);
            if (dev_ && tsp_timer_start(dev_) != TSP_OK) {
                release(dev_, s);
                dev_ = nullptr;
            }
        }
        host_start_ = Clock::now();
    }
    ScopedTimer(const ScopedTimer &) = delete;
    ScopedTimer &operator=(const ScopedTimer &) = delete;

    ~ScopedTimer() {
        if (!active_) return;
        double run_us = std::chrono::duration<double, std::micro>(Clock::now() - host_start_).count();
        if (dev_) {
            double ms = 0.0;
            if (tsp_timer_stop(dev_, &ms) == TSP_OK) run_us = ms * 1e3;
            release(dev_, dev_stream_);
        }
        Stats &s = table()[name_];
        s.total_us += run_us;
        s.sum_sq_us += run_us * run_us;
        ++s.count;
        std::cout << "[Timer: " << name_ << "] Run: " << format_us(run_us) << " | Total: " << format_us(s.total_us)
                  << " | Avg: " << format_us(s.mean_us()) << " | StdDev: " << format_us(s.stdev_us()) << " ("
                  << s.count << " runs)\n";
    }

    static void printOverview() {
        std::cout << "--- --- --- ScopedTimer Profiling Overview --- --- ---" << std::endl;
        std::vector<std::pair<std::string, Stats>> rows(table().begin(), table().end());
        std::sort(rows.begin(), rows.end(), [](const auto &a, const auto &b) { return a.second.total_us > b.second.total_us; });
        for (const auto &[name, s] : rows) {
            std::cout << "[Timer: " << std::left << std::setw(35) << name << "] " << std::right
                      << "Executions: " << std::setw(6) << s.count << " | Total: " << std::setw(8)
                      << format_us(s.total_us) << " | Avg: " << std::setw(8) << format_us(s.mean_us())
                      << " | StdDev: " << std::setw(9) << format_us(s.stdev_us()) << std::endl;
        }
    }

    static void reset() { table().clear(); }

    inline static bool enabled = false;

    // Additions without a reference counterpart: the thread's timing stream
    // (a hipStream_t; nullptr: the host clock).
    static void *timing_stream() { return current(); }
    static void set_timing_stream(void *hip_stream) { current() = hip_stream; }
    // Names the timing stream for the scopes opened during its lifetime (a
    // nullptr stream leaves the current one).
    class StreamGuard {
      public:
        explicit StreamGuard(void *hip_stream) : prev_(current()) {
            if (hip_stream) current() = hip_stream;
        }
        ~StreamGuard() { current() = prev_; }
        StreamGuard(const StreamGuard &) = delete;
        StreamGuard &operator=(const StreamGuard &) = delete;

      private:
        void *prev_;
    };

  private:
    static void *&current() {
        thread_local void *s = nullptr;
        return s;
    }
    // event pairs are reused per stream; the pool lives until the process ends
    static std::vector<std::pair<void *, tsp_timer *>> &pool() {
        thread_local std::vector<std::pair<void *, tsp_timer *>> *p = new std::vector<std::pair<void *, tsp_timer *>>;
        return *p;
    }
    static tsp_timer *acquire(void *stream) {
        auto &p = pool();
        for (std::size_t i = 0; i < p.size(); ++i)
            if (p[i].first == stream) {
                tsp_timer *t = p[i].second;
                p[i] = p.back();
                p.pop_back();
                return t;
            }
        tsp_timer *t = nullptr;
        return tsp_timer_create(stream, &t) == TSP_OK ? t : nullptr;
    }
    static void release(tsp_timer *t, void *stream) { pool().emplace_back(stream, t); }
    static std::unordered_map<std::string, Stats> &table() {
        static std::unordered_map<std::string, Stats> t;
        return t;
    }
    static std::string format_us(double us) {
        std::ostringstream o;
        o << std::fixed << std::setprecision(1);
        if (us < 1e3) o << us << " µs";
        else if (us < 1e6) o << us / 1e3 << " ms";
        else o << us / 1e6 << " s";
        return o.str();
    }

    std::string name_;
    bool active_ = false;
    tsp_timer *dev_ = nullptr;
    void *dev_stream_ = nullptr;
    Clock::time_point host_start_{};
};

} // namespace treesap
