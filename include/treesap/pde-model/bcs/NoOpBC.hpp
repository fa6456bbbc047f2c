// NoOpBC.hpp -- physical-boundary functor that leaves ghost cells untouched
// (pde-model/bcs/NoOpBC.hpp:22-47). On the device this is TSP_BC_NOOP: ghost
// cells outside a non-periodic face keep whatever the register holds (zero
// from allocation, or the value last written there).
#pragma once

#include "../PDEModelConcepts.hpp"

namespace treesap {

template <int ncomp_, int ngrow_> struct NoOpBC {
    static constexpr int ncomp = ncomp_;
    static constexpr int ngrow = ngrow_;

    AMREX_GPU_HOST_DEVICE void operator()(const amrex::IntVect &, const amrex::Array4<amrex::Real> &, int, int, const amrex::GeometryData &,
                    amrex::Real, const amrex::BCRec *, int, int) const {}
};

static_assert(BCConcept<NoOpBC<1, 1>>);

} // namespace treesap
