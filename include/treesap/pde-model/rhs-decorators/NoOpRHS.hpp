// NoOpRHS.hpp -- right-hand side that is zero for every component
// (pde-model/rhs-decorators/NoOpRHS.hpp:19-33); used in concept checks.
#pragma once

#include "../PDEModelConcepts.hpp"

namespace treesap {

template <int ncomp_, int ngrow_> struct NoOpRHS {
    static constexpr int ncomp = ncomp_;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect, amrex::Real, const amrex::Array4<const amrex::Real> &,
                           const amrex::GeometryData &, Workspace &) const {
        return 0.0;
    }
};

static_assert(RHSConcept<NoOpRHS<1, 1>>);

} // namespace treesap
