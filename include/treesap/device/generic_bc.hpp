// generic_bc.hpp -- device application of a BCConcept functor the library does
// not carry (anything but NoOpBC and DirichletBC): the physical-boundary step
// of SingleLevelIntegrator::FillPatch (SingleLevelIntegrator.hpp:153-163),
// where amrex::PhysBCFunct<amrex::GpuBndryFuncFab<BCType>>(model.get_bc())
// calls the functor on the ghost cells outside the domain, compiled from the
// user's functor in the user's translation unit.
//
// The library fills periodic images itself and then calls `bc_callback`
// (tsp_set_bc_callback) with the register just filled and the fill's time.
// The callback launches generic_bc_kernel on the solver's stream: one thread
// per ghost cell of the shell (z slabs, then y slabs inside the valid z range,
// then x slabs inside the valid y, z range, as the library's own fill kernel
// enumerates them). A cell outside the domain along a non-periodic axis calls
//   bc(iv, dest, 0, ncomp, geom, time, bcr, 0, 0)
// with iv in the domain's own indices, dest an Array4 view of the whole
// register (ghosts included) and bcr the model's BCRec per component. All
// such cells run concurrently: the functor may read any valid cell and its own
// ghost cell, and writes its own ghost cell (as in AMReX, whose boundary
// boxes' cells also run in parallel).
//
// The kernel exists only where the including translation unit is compiled by a
// HIP compiler; a host-only build that steps such a model fails to compile
// with a message saying so.
#pragma once

#include "../amrex_types.hpp"
#include "solver.hpp"
#include "treesap_hip.h"

#include <cstdint>
#include <stdexcept>

namespace treesap::device {

// what the boundary callback needs: a copy of the functor, the geometry and
// the model's boundary records
template <class BC> struct GenericBC {
    BC bc;
    amrex::GeometryData geom;
    amrex::BCRec bcr[TSP_MAX_COMP];
};

template <class BC> int bc_callback(void *user, const tsp_bc_view *v, void *stream);

template <class> constexpr bool generic_bc_needs_hip = false;

template <class BC> class BCRunner {
  public:
    BCRunner(const BC &bc, const amrex::GeometryData &g, const amrex::Vector<amrex::BCRec> &recs) : st_{bc, g, {}} {
        if (recs.size() > (std::size_t)TSP_MAX_COMP)
            throw std::invalid_argument("treesap: more boundary records than components");
        for (std::size_t c = 0; c < recs.size(); ++c) st_.bcr[c] = recs[c];
    }
    // hand the callback (and this runner's address) to a single-box solver;
    // done before every stepping call, so a moved integrator passes its own copy
    void attach(Solver &s) {
#if defined(__HIP__)
        if (!s.handle()) throw std::invalid_argument("treesap: a custom boundary functor runs on a single-box solver");
        check(tsp_set_bc_callback(s.handle(), &bc_callback<BC>, &st_), "tsp_set_bc_callback");
#else
        (void)s;
        static_assert(generic_bc_needs_hip<BC>, "treesap: this boundary functor is not one the device library carries "
                                                "(NoOpBC, DirichletBC); its kernel is compiled from the functor, so "
                                                "compile this program with hipcc");
#endif
    }

  private:
    GenericBC<BC> st_;
};

} // namespace treesap::device

#if defined(__HIP__)
#include <hip/hip_runtime.h>

namespace treesap::device {

template <class BC>
__global__ __launch_bounds__(256) void generic_bc_kernel(const GenericBC<BC> g, const tsp_bc_view v, long long total) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int G = v.ghost, nx = v.n[0], ny = v.n[1], nz = v.n[2];
    const int gx = nx + 2 * G, gy = ny + 2 * G;
    const long long na = 2LL * G * gy * gx, nb = (long long)nz * 2 * G * gx;
    int i, j, k;
    if (t < na) { // z slabs
        const long long plane = (long long)gy * gx, r = t % plane;
        const int kk = (int)(t / plane);
        j = (int)(r / gx) - G;
        i = (int)(r % gx) - G;
        k = kk < G ? kk - G : nz + (kk - G);
    } else if (t < na + nb) { // y slabs inside the valid z range
        const long long u = t - na, plane = 2LL * G * gx, r = u % plane;
        k = (int)(u / plane);
        const int jj = (int)(r / gx);
        i = (int)(r % gx) - G;
        j = jj < G ? jj - G : ny + (jj - G);
    } else { // x slabs inside the valid y, z range
        const long long u = t - na - nb, plane = (long long)ny * 2 * G, r = u % plane;
        k = (int)(u / plane);
        j = (int)(r / (2 * G));
        const int ii = (int)(r % (2 * G));
        i = ii < G ? ii - G : nx + (ii - G);
    }
    const bool phys = ((i < 0 || i >= nx) && !v.periodic[0]) || ((j < 0 || j >= ny) && !v.periodic[1]) ||
                      ((k < 0 || k >= nz) && !v.periodic[2]);
    if (!phys) return;
    const amrex::IntVect lo = g.geom.Domain().smallEnd();
    double *p = v.reg - (G + G * v.sj + G * v.sk);
    const amrex::Array4<amrex::Real> dest(p, v.sj, v.sk, v.sc, amrex::Dim3{lo[0] - G, lo[1] - G, lo[2] - G},
                                          amrex::Dim3{lo[0] + nx + G, lo[1] + ny + G, lo[2] + nz + G}, v.ncomp);
    const amrex::IntVect iv{lo[0] + i, lo[1] + j, lo[2] + k};
    g.bc(iv, dest, 0, v.ncomp, g.geom, v.time, g.bcr, 0, 0);
}

// tsp_bc_fn: launch the boundary kernel of `user` (a GenericBC<BC>) on the solver's
// stream; returns 0, or 2 for a failed launch
template <class BC> int bc_callback(void *user, const tsp_bc_view *v, void *stream) {
    const auto *g = static_cast<const GenericBC<BC> *>(user);
    const int G = v->ghost;
    if (G == 0) return 0;
    const long long gx = v->n[0] + 2 * G, gy = v->n[1] + 2 * G;
    const long long total =
        2LL * G * gy * gx + (long long)v->n[2] * 2 * G * gx + (long long)v->n[2] * v->n[1] * 2 * G;
    hipLaunchKernelGGL((generic_bc_kernel<BC>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, *g, *v, total);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

} // namespace treesap::device
#endif
