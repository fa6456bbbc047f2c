// solver.hpp -- C++ ownership wrapper around one tsp_solver of the C ABI
// (include/treesap_hip.h). Failures become std::runtime_error carrying the
// library's message; nothing here touches device memory directly.
#pragma once

#include "treesap_hip.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace treesap::device {

inline void check(tsp_status st, const char *what) {
    if (st != TSP_OK) throw std::runtime_error(std::string("treesap: ") + what + " failed: " + tsp_last_error());
}

class Solver {
  public:
    Solver() = default;
    explicit Solver(const tsp_desc &d, void *hip_stream = nullptr) : desc_(d) {
        check(tsp_create(&desc_, hip_stream, &s_), "tsp_create");
    }
    ~Solver() {
        if (s_) tsp_destroy(s_);
    }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;
    Solver(Solver &&o) noexcept : s_(std::exchange(o.s_, nullptr)), desc_(o.desc_) {}
    Solver &operator=(Solver &&o) noexcept {
        if (this != &o) {
            if (s_) tsp_destroy(s_);
            s_ = std::exchange(o.s_, nullptr);
            desc_ = o.desc_;
        }
        return *this;
    }

    // valid region, a[((c * nz + k) * ny + j) * nx + i]
    void set_state(const double *host) { check(tsp_set_state(s_, host, 0), "tsp_set_state"); }
    void get_state(double *host) const { check(tsp_get_state(s_, host, 0), "tsp_get_state"); }
    void rk4_steps(std::int64_t n) { check(tsp_rk4_steps(s_, n), "tsp_rk4_steps"); }
    void graph_rk4_steps(std::int64_t n) { check(tsp_graph_rk4_steps(s_, n), "tsp_graph_rk4_steps"); }
    void euler_steps(std::int64_t n) { check(tsp_euler_steps(s_, n), "tsp_euler_steps"); }
    void sync() { check(tsp_sync(s_), "tsp_sync"); }
    tsp_kernel_cfg autotune(int steps) {
        tsp_kernel_cfg c{};
        check(tsp_autotune(s_, steps, &c), "tsp_autotune");
        return c;
    }
    tsp_kernel_cfg kernel() const {
        tsp_kernel_cfg c{};
        check(tsp_get_kernel(s_, &c), "tsp_get_kernel");
        return c;
    }
    void set_kernel(const tsp_kernel_cfg &c) { check(tsp_set_kernel(s_, &c), "tsp_set_kernel"); }
    std::pair<double, std::int64_t> time_and_steps() const {
        double t = 0.0;
        std::int64_t n = 0;
        check(tsp_get_time(s_, &t, &n), "tsp_get_time");
        return {t, n};
    }

    const tsp_desc &desc() const { return desc_; }
    std::int64_t cells() const { return (std::int64_t)desc_.n[0] * desc_.n[1] * desc_.n[2]; }
    tsp_solver *handle() const { return s_; }

  private:
    tsp_solver *s_ = nullptr;
    tsp_desc desc_{};
};

} // namespace treesap::device
