// solver.hpp -- C++ ownership wrapper around one device solve of the C ABI
// (include/treesap_hip.h): a single-box tsp_solver, or a level of several
// boxes (tsp_level_*, a tsp_batch) when the block size cuts the domain.
// Failures become std::runtime_error carrying the library's message; nothing
// here touches device memory directly.
#pragma once

#include "treesap_hip.h"

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace treesap::device {

inline void check(tsp_status st, const char *what) {
    if (st != TSP_OK) throw std::runtime_error(std::string("treesap: ") + what + " failed: " + tsp_last_error());
}

class Solver {
  public:
    Solver() = default;
    // `block`: the largest box per axis (SingleLevelIntegrator(model, block_size)); a
    // block smaller than the domain along any axis splits the level into boxes
    // (the (phi, Pi), (phi, u) and NonLinear systems; the tensor system,
    // functors on the generic kernel and custom boundary functors run one
    // box, results being the same)
    explicit Solver(const tsp_desc &d, const int *block = nullptr, void *hip_stream = nullptr) : desc_(d) {
        bool cut = false;
        if (block)
            for (int a = 0; a < 3; ++a) cut = cut || block[a] < d.n[a];
        const bool level_system = (d.system == TSP_SYS_WAVE_PHIPI || d.system == TSP_SYS_WAVE_PHIU ||
                                   d.system == TSP_SYS_NONLINEAR) &&
                                  d.bc != TSP_BC_CUSTOM;
        if (cut && level_system) {
            const int32_t blk[3] = {block[0], block[1], block[2]};
            check(tsp_level_create(&desc_, blk, hip_stream, &l_), "tsp_level_create");
        } else {
            check(tsp_create(&desc_, hip_stream, &s_), "tsp_create");
        }
    }
    ~Solver() { release(); }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;
    Solver(Solver &&o) noexcept : s_(std::exchange(o.s_, nullptr)), l_(std::exchange(o.l_, nullptr)), desc_(o.desc_) {}
    Solver &operator=(Solver &&o) noexcept {
        if (this != &o) {
            release();
            s_ = std::exchange(o.s_, nullptr);
            l_ = std::exchange(o.l_, nullptr);
            desc_ = o.desc_;
        }
        return *this;
    }

    // valid region, a[((c * nz + k) * ny + j) * nx + i]
    void set_state(const double *host) {
        if (l_) check(tsp_level_set_state(l_, host, 0), "tsp_level_set_state");
        else check(tsp_set_state(s_, host, 0), "tsp_set_state");
    }
    // initialise_impl on the device: postfix expression programs (FunctionInit::program)
    // evaluated at the cell centres; index_lo is the domain's first cell index
    void init_expr(const tsp_expr_op *ops, const std::int32_t *start, const std::int32_t index_lo[3]) {
        if (l_) check(tsp_level_init_expr(l_, ops, start, index_lo), "tsp_level_init_expr");
        else check(tsp_init_expr(s_, ops, start, index_lo), "tsp_init_expr");
    }
    void get_state(double *host) const {
        if (l_) check(tsp_level_get_state(l_, host, 0), "tsp_level_get_state");
        else check(tsp_get_state(s_, host, 0), "tsp_get_state");
    }
    void rk4_steps(std::int64_t n) {
        if (l_) check(tsp_batch_rk4_steps(l_, n), "tsp_batch_rk4_steps");
        else check(tsp_rk4_steps(s_, n), "tsp_rk4_steps");
    }
    void graph_rk4_steps(std::int64_t n) {
        if (l_) check(tsp_batch_graph_rk4_steps(l_, n), "tsp_batch_graph_rk4_steps");
        else check(tsp_graph_rk4_steps(s_, n), "tsp_graph_rk4_steps");
    }
    void euler_steps(std::int64_t n) {
        if (l_) check(tsp_batch_euler_steps(l_, n), "tsp_batch_euler_steps");
        else check(tsp_euler_steps(s_, n), "tsp_euler_steps");
    }
    void sync() {
        if (l_) check(tsp_batch_sync(l_), "tsp_batch_sync");
        else check(tsp_sync(s_), "tsp_sync");
    }
    // stage-kernel configuration chosen by the library's autotuner (for a level
    // the batch configuration is tuned and kept; the returned single-box
    // configuration is then zero)
    tsp_kernel_cfg autotune(int steps) {
        tsp_kernel_cfg c{};
        if (l_) {
            tsp_batch_cfg bc{};
            check(tsp_batch_autotune(l_, steps, &bc), "tsp_batch_autotune");
        } else {
            check(tsp_autotune(s_, steps, &c), "tsp_autotune");
        }
        return c;
    }
    tsp_kernel_cfg kernel() const {
        tsp_kernel_cfg c{};
        if (s_) check(tsp_get_kernel(s_, &c), "tsp_get_kernel");
        return c;
    }
    void set_kernel(const tsp_kernel_cfg &c) {
        if (!s_) throw std::runtime_error("treesap: a level's boxes use the batch configuration");
        check(tsp_set_kernel(s_, &c), "tsp_set_kernel");
    }
    std::pair<double, std::int64_t> time_and_steps() const {
        double t = 0.0;
        std::int64_t n = 0;
        if (l_) check(tsp_batch_get_time(l_, 0, &t, &n), "tsp_batch_get_time");
        else check(tsp_get_time(s_, &t, &n), "tsp_get_time");
        return {t, n};
    }

    const tsp_desc &desc() const { return desc_; }
    std::int64_t cells() const { return (std::int64_t)desc_.n[0] * desc_.n[1] * desc_.n[2]; }
    bool is_level() const { return l_ != nullptr; }
    int boxes() const {
        if (!l_) return 1;
        int32_t n = 0;
        check(tsp_batch_size(l_, &n, nullptr), "tsp_batch_size");
        return n;
    }
    // the hipStream_t the solve is enqueued on
    void *stream() const {
        void *st = nullptr;
        if (l_) check(tsp_batch_get_stream(l_, &st), "tsp_batch_get_stream");
        else if (s_) check(tsp_get_stream(s_, &st), "tsp_get_stream");
        return st;
    }
    // every box's lo and hi (inclusive, relative to the domain's first cell), in level order
    std::vector<std::array<int, 6>> box_bounds() const {
        const int nb = boxes();
        std::vector<std::array<int, 6>> out(nb);
        if (!l_) {
            out[0] = {0, 0, 0, desc_.n[0] - 1, desc_.n[1] - 1, desc_.n[2] - 1};
            return out;
        }
        std::vector<int32_t> lh(6 * (std::size_t)nb);
        check(tsp_level_boxes(l_, lh.data()), "tsp_level_boxes");
        for (int b = 0; b < nb; ++b)
            for (int q = 0; q < 6; ++q) out[b][q] = lh[6 * b + q];
        return out;
    }
    tsp_solver *handle() const { return s_; }   // single box (null for a level)
    tsp_batch *level() const { return l_; }     // a level (null for a single box)

  private:
    void release() {
        if (s_) tsp_destroy(s_);
        if (l_) tsp_batch_destroy(l_);
        s_ = nullptr;
        l_ = nullptr;
    }
    tsp_solver *s_ = nullptr;
    tsp_batch *l_ = nullptr;
    tsp_desc desc_{};
};

} // namespace treesap::device
