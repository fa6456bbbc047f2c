// ScopedTimer.hpp -- wall-clock timer for a named scope with aggregated
// statistics (util/ScopedTimer.hpp:34-149): when ScopedTimer::enabled is true
// each timer prints its run time and running totals when it goes out of
// scope, and printOverview() lists every name by total time. Device work is
// asynchronous, so scopes that contain device steps measure launch time only
// unless the integrator synchronises inside them (it does when timers are on).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
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
        Clock::duration total{0};
        double sum_sq_us = 0.0;
        std::size_t count = 0;
        double mean_us() const { return std::chrono::duration<double, std::micro>(total).count() / count; }
        double stdev_us() const {
            const double m = mean_us();
            const double var = sum_sq_us / count - m * m;
            return var > 0.0 ? std::sqrt(var) : 0.0;
        }
    };

  public:
    explicit ScopedTimer(std::string name) : name_(std::move(name)), start_(Clock::now()) {}
    ScopedTimer(const ScopedTimer &) = delete;
    ScopedTimer &operator=(const ScopedTimer &) = delete;

    ~ScopedTimer() {
        if (!enabled) return;
        const Clock::duration run = Clock::now() - start_;
        const double run_us = std::chrono::duration<double, std::micro>(run).count();
        Stats &s = table()[name_];
        s.total += run;
        s.sum_sq_us += run_us * run_us;
        ++s.count;
        std::cout << "[Timer: " << name_ << "] Run: " << format_us(run_us) << " | Total: "
                  << format_us(std::chrono::duration<double, std::micro>(s.total).count())
                  << " | Avg: " << format_us(s.mean_us()) << " | StdDev: " << format_us(s.stdev_us()) << " ("
                  << s.count << " runs)\n";
    }

    static void printOverview() {
        std::cout << "--- --- --- ScopedTimer Profiling Overview --- --- ---" << std::endl;
        std::vector<std::pair<std::string, Stats>> rows(table().begin(), table().end());
        std::sort(rows.begin(), rows.end(), [](const auto &a, const auto &b) { return a.second.total > b.second.total; });
        for (const auto &[name, s] : rows) {
            std::cout << "[Timer: " << std::left << std::setw(35) << name << "] " << std::right
                      << "Executions: " << std::setw(6) << s.count << " | Total: " << std::setw(8)
                      << format_us(std::chrono::duration<double, std::micro>(s.total).count()) << " | Avg: "
                      << std::setw(8) << format_us(s.mean_us()) << " | StdDev: " << std::setw(9)
                      << format_us(s.stdev_us()) << std::endl;
        }
    }

    static void reset() { table().clear(); }

    inline static bool enabled = false;

  private:
    static std::unordered_map<std::string, Stats> &table() {
        static std::unordered_map<std::string, Stats> t;
        return t;
    }
    static std::string format_us(double us) {
        std::ostringstream o;
        o << std::fixed << std::setprecision(1);
        if (us < 1e3) o << us << " us";
        else if (us < 1e6) o << us / 1e3 << " ms";
        else o << us / 1e6 << " s";
        return o.str();
    }

    std::string name_;
    Clock::time_point start_;
};

} // namespace treesap
