// generic.hpp -- device evaluation of any RHSConcept functor: the fused RK4 /
// Euler stage of integrators/SingleLevelRK4Integrator.hpp:110-157 (runStage)
// and SingleLevelEulerIntegrator.hpp:66-88, compiled from the user's functor
// in the user's translation unit, as the reference compiles it into its
// ParallelFor lambda.
//
// The library (tsp_rk4_steps_custom / tsp_euler_steps_custom) keeps the
// registers, fills their ghosts before every stage and rotates the roles; it
// calls `stage_callback` once per stage, which launches generic_stage_kernel
// on the solver's stream. One thread per valid cell, 64 x 4 cells of one z
// plane per workgroup (x fastest, so each wave's loads of a register row are
// contiguous; no index is divided); per cell one Workspace and the components
// in order 0..ncomp-1, each evaluated on an Array4 view of the ghost-filled
// source register. All components are evaluated before any is stored: the
// functor reads only the source register, which a stage never writes, so the
// values are those of storing each at once, and the loads the components'
// stencils share are issued once (1.3x faster on a (phi, Pi) functor). The
// stores (non-temporal, 4 % faster with the centre reads below) follow the
// reference's expressions:
//   stage 0:  k = F<c>(init);  acc = init + b k;  arg_next = init + a_next k
//   stage s:  k = F<c>(arg);   acc += b k;        arg_next = init + a_next k (s < 3)
//   Euler:    new = old + dt F<c>(old)
// Registered systems keep the library's tiled kernels; this path runs the
// functors the library does not carry (any stencil order, D3, antisymmetric
// parts, Levi-Civita contractions, position- or time-dependent terms).
//
// Device code: include this header only in a translation unit compiled by a
// HIP compiler (hipcc --offload-arch=gfx950).
#pragma once

#include "../amrex_types.hpp"
#include "../util/static_for.hpp"
#include "generic_runner.hpp"
#include "treesap_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace treesap::device {

// Array4 view of a register of the stage view (ghosts included), indexed as
// the domain is: its first valid cell is `lo`, the domain's smallEnd
template <class T>
__host__ __device__ inline amrex::Array4<T> register_view(T *origin, const tsp_stage_view &v, const amrex::IntVect &lo) {
    const int G = v.ghost;
    T *p = origin - (G + G * v.sj + G * v.sk);
    return amrex::Array4<T>(p, v.sj, v.sk, v.sc, amrex::Dim3{lo[0] - G, lo[1] - G, lo[2] - G},
                            amrex::Dim3{lo[0] + v.n[0] + G, lo[1] + v.n[1] + G, lo[2] + v.n[2] + G}, v.ncomp);
}

template <class RHS, int KIND> // KIND: 0 first RK4 stage, 1 middle, 2 last, 3 Euler
__global__ __launch_bounds__(256) void generic_stage_kernel(const RHS rhs, const amrex::GeometryData geom,
                                                            const tsp_stage_view v) {
    constexpr int ncomp = RHS::ncomp;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= v.n[0] || j >= v.n[1]) return;
    // cells carry the domain's own indices, as the reference's ParallelFor over validbox gives them
    const amrex::IntVect lo = geom.Domain().smallEnd();
    const amrex::Array4<const amrex::Real> src = register_view<const amrex::Real>(v.src, v, lo);
    const amrex::Array4<const amrex::Real> init = register_view<const amrex::Real>(v.init, v, lo);
    const amrex::Array4<amrex::Real> acc = register_view<amrex::Real>(v.acc, v, lo);
    const amrex::Real b = v.b, a_next = v.a_next, t = v.time;
    const amrex::IntVect idx{lo[0] + i, lo[1] + j, lo[2] + k};
    typename RHS::Workspace workspace{};
    amrex::Real kv[ncomp];
    treesap::static_for<ncomp>([&]<int comp>() { kv[comp] = rhs.template operator()<comp>(idx, t, src, geom, workspace); });
    treesap::static_for<ncomp>([&]<int comp>() {
        // Stores, and the centre reads of registers the stencils do not read (init in
        // stages 1 and 2, acc), are streamed (non-temporal) past the caches the
        // stencil reads of src use; the arithmetic is unchanged.
        if constexpr (KIND == 3) {
            __builtin_nontemporal_store(init(idx, comp) + b * kv[comp], &acc(idx, comp));
        } else {
            const amrex::Real iv = KIND == 1 ? __builtin_nontemporal_load(&init(idx, comp)) : init(idx, comp);
            if constexpr (KIND == 0) __builtin_nontemporal_store(iv + b * kv[comp], &acc(idx, comp));
            else __builtin_nontemporal_store(__builtin_nontemporal_load(&acc(idx, comp)) + b * kv[comp], &acc(idx, comp));
            if constexpr (KIND != 2) {
                const amrex::Array4<amrex::Real> argn = register_view<amrex::Real>(v.argn, v, lo);
                __builtin_nontemporal_store(iv + a_next * kv[comp], &argn(idx, comp));
            }
        }
    });
}

// tsp_stage_fn: launch the stage kernel of `user` (a GenericStage<RHS>) on the solver's stream;
// returns 0, or 1 for an unknown stage, 2 for a failed launch, 3 for a box whose y or z
// extent exceeds the launch grid (more than 4 * 65535 cells along y or 65535 along z)
template <class RHS> int stage_callback(void *user, const tsp_stage_view *v, void *stream) {
    const auto *g = static_cast<const GenericStage<RHS> *>(user);
    if ((v->n[1] + 3) / 4 > 65535 || v->n[2] > 65535) return 3; // beyond the grid's y and z extents
    const dim3 blocks((v->n[0] + 63) / 64, (v->n[1] + 3) / 4, v->n[2]), threads(64, 4);
    const hipStream_t st = (hipStream_t)stream;
    switch (v->stage) {
    case 0: hipLaunchKernelGGL((generic_stage_kernel<RHS, 0>), blocks, threads, 0, st, g->rhs, g->geom, *v); break;
    case 1:
    case 2: hipLaunchKernelGGL((generic_stage_kernel<RHS, 1>), blocks, threads, 0, st, g->rhs, g->geom, *v); break;
    case 3: hipLaunchKernelGGL((generic_stage_kernel<RHS, 2>), blocks, threads, 0, st, g->rhs, g->geom, *v); break;
    case 4: hipLaunchKernelGGL((generic_stage_kernel<RHS, 3>), blocks, threads, 0, st, g->rhs, g->geom, *v); break;
    default: return 1;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

} // namespace treesap::device
