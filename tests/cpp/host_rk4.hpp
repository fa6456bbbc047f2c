// host_rk4.hpp -- test helper: the reference's RK4 / Euler stepping of any
// RHSConcept functor on the host, for comparison with the device's generic
// kernel. Registers with G ghost layers; before every stage the source
// register's ghosts are filled as FillPatch does on one box (a periodic axis
// takes the periodic image, a ghost outside a non-periodic face takes the
// Dirichlet values, or keeps its memory with NoOp, or, with a custom boundary
// functor set, takes what the functor writes after all periodic images are
// in place, called once per cell with the fill's time), then
// integrators/SingleLevelRK4Integrator.hpp:110-157 (runStage) or
// SingleLevelEulerIntegrator.hpp:66-88 run cell by cell, components in order
// with one Workspace per cell. Cells carry the domain's own indices (its
// smallEnd need not be 0).
#pragma once

#include "amrex_types.hpp"
#include "util/static_for.hpp"

#include <array>
#include <cstdint>
#include <functional>
#include <utility>
#include <vector>

template <class RHS> struct HostStepper {
    static constexpr int nc = RHS::ncomp;
    RHS rhs;
    amrex::GeometryData geom;
    int n[3], lo[3], G;
    bool dirichlet;
    std::array<double, 12> target{};
    double dt, time = 0.0;
    std::vector<double> reg[4]; // old, new, arg, arg_next
    int role[4] = {0, 1, 2, 3};
    // a custom boundary functor: (cell in domain indices, register view, fill time)
    std::function<void(const amrex::IntVect &, const amrex::Array4<double> &, double)> custom_bc;

    HostStepper(const RHS &r, const amrex::Geometry &g, double dt_, bool dirichlet_, const double *tgt)
        : rhs(r), geom(g.data()), G(RHS::ngrow), dirichlet(dirichlet_), dt(dt_) {
        for (int a = 0; a < 3; ++a) {
            n[a] = g.Domain().length(a);
            lo[a] = g.Domain().smallEnd(a);
        }
        for (int c = 0; c < nc; ++c) target[c] = tgt ? tgt[c] : 0.0;
        for (auto &v : reg) v.assign(size(), 0.0);
    }
    std::size_t size() const { return (std::size_t)(n[0] + 2 * G) * (n[1] + 2 * G) * (n[2] + 2 * G) * nc; }
    template <class T> amrex::Array4<T> view(int r) {
        return amrex::Array4<T>(reinterpret_cast<T *>(reg[r].data()), {lo[0] - G, lo[1] - G, lo[2] - G},
                                {lo[0] + n[0] + G, lo[1] + n[1] + G, lo[2] + n[2] + G}, nc);
    }
    void set_state(const std::vector<double> &u) { // valid layout a[((c*nz+k)*ny+j)*nx+i]
        auto a = view<double>(role[0]);
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < n[2]; ++k)
                for (int j = 0; j < n[1]; ++j)
                    for (int i = 0; i < n[0]; ++i)
                        a(lo[0] + i, lo[1] + j, lo[2] + k, c) = u[(((std::size_t)c * n[2] + k) * n[1] + j) * n[0] + i];
    }
    std::vector<double> get_state() {
        std::vector<double> u((std::size_t)nc * n[0] * n[1] * n[2]);
        auto a = view<double>(role[0]);
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < n[2]; ++k)
                for (int j = 0; j < n[1]; ++j)
                    for (int i = 0; i < n[0]; ++i)
                        u[(((std::size_t)c * n[2] + k) * n[1] + j) * n[0] + i] = a(lo[0] + i, lo[1] + j, lo[2] + k, c);
        return u;
    }
    void fill(int r, double t) {
        auto a = view<double>(r);
        for (int pass = 0; pass < 2; ++pass) // periodic images, then the physical boundary
            for (int k = -G; k < n[2] + G; ++k)
                for (int j = -G; j < n[1] + G; ++j)
                    for (int i = -G; i < n[0] + G; ++i) {
                        const int iv[3] = {i, j, k};
                        bool inside = true, phys = false;
                        int w[3];
                        for (int d = 0; d < 3; ++d) {
                            const bool out = iv[d] < 0 || iv[d] >= n[d];
                            inside = inside && !out;
                            if (out && !geom.is_periodic[d]) phys = true;
                            w[d] = ((iv[d] % n[d]) + n[d]) % n[d];
                        }
                        if (inside || phys != (pass == 1)) continue;
                        if (phys && custom_bc) {
                            custom_bc(amrex::IntVect{lo[0] + i, lo[1] + j, lo[2] + k}, a, t);
                            continue;
                        }
                        for (int c = 0; c < nc; ++c) {
                            if (phys) {
                                if (dirichlet) a(lo[0] + i, lo[1] + j, lo[2] + k, c) = target[c];
                            } else {
                                a(lo[0] + i, lo[1] + j, lo[2] + k, c) = a(lo[0] + w[0], lo[1] + w[1], lo[2] + w[2], c);
                            }
                        }
                    }
    }
    void rk4_step() {
        static constexpr double af[4] = {0.0, 0.5, 0.5, 1.0};
        static constexpr double bf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
        int &OLD = role[0], &NEW = role[1], &ARG = role[2], &ARGN = role[3];
        fill(OLD, time);
        for (int s = 0; s < 4; ++s) {
            const double a_next = dt * af[(s + 1) % 4];
            const double t = time + dt * af[s];
            const double b = dt * bf[s];
            if (s > 0) {
                std::swap(ARG, ARGN);
                fill(ARG, t);
            }
            const auto init = view<const double>(OLD);
            const auto src = view<const double>(s == 0 ? OLD : ARG);
            auto acc = view<double>(NEW);
            auto argn = view<double>(ARGN);
            for (int k = 0; k < n[2]; ++k)
                for (int j = 0; j < n[1]; ++j)
                    for (int i = 0; i < n[0]; ++i) {
                        const amrex::IntVect idx{lo[0] + i, lo[1] + j, lo[2] + k};
                        typename RHS::Workspace ws{};
                        treesap::static_for<nc>([&]<int c>() {
                            const double stage = rhs.template operator()<c>(idx, t, src, geom, ws);
                            if (s == 0) acc(idx, c) = init(idx, c) + b * stage;
                            else acc(idx, c) += b * stage;
                            if (s < 3) argn(idx, c) = init(idx, c) + a_next * stage;
                        });
                    }
        }
        std::swap(OLD, NEW);
        time += dt;
    }
    void euler_step() {
        int &OLD = role[0], &NEW = role[1];
        fill(OLD, time);
        const auto old = view<const double>(OLD);
        auto nw = view<double>(NEW);
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) {
                    const amrex::IntVect idx{lo[0] + i, lo[1] + j, lo[2] + k};
                    typename RHS::Workspace ws{};
                    treesap::static_for<nc>([&]<int c>() {
                        nw(idx, c) = old(idx, c) + dt * rhs.template operator()<c>(idx, time, old, geom, ws);
                    });
                }
        std::swap(OLD, NEW);
        time += dt;
    }
};
