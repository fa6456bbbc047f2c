// DirichletBC.hpp -- physical-boundary functor writing fixed values into every
// ghost cell outside the domain box (pde-model/bcs/DirichletBC.hpp:21-66):
// all ncomp components are set, whatever dcomp/numcomp/bcr say. On the device
// this is TSP_BC_DIRICHLET with target[] = boundary_values.
#pragma once

#include "../PDEModelConcepts.hpp"

namespace treesap {

template <int ncomp_, int ngrow_> struct DirichletBC {
    static constexpr int ncomp = ncomp_;
    static constexpr int ngrow = ngrow_;

    amrex::GpuArray<amrex::Real, ncomp> boundary_values{};

    AMREX_GPU_HOST_DEVICE void operator()(const amrex::IntVect &iv, const amrex::Array4<amrex::Real> &dest, int, int,
                    const amrex::GeometryData &geom, amrex::Real, const amrex::BCRec *, int, int) const {
        if (geom.Domain().contains(iv)) return;
        for (int c = 0; c < ncomp; ++c) dest(iv, c) = boundary_values[c];
    }
};

static_assert(BCConcept<DirichletBC<1, 1>>);

} // namespace treesap
