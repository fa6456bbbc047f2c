// KODecorator.hpp -- adds Kreiss-Oliger dissipation to a right-hand side
// (pde-model/rhs-decorators/KODecorator.hpp:32-74):
//   F_c + strength * (((0 + KO_x u_c) + KO_y u_c) + KO_z u_c)
// with KO of order r (default RHS::ngrow, 0 < r <= ngrow). The reference's
// call operator passes its arguments to kreiss_oliger in a different order
// than kreiss_oliger declares them (KODecorator.hpp:65-67 against
// util/stencil.hpp:77-80); this one passes (arr, comp, idx, geom), the KO
// stencil of component comp at idx, which is what the decorator documents.
#pragma once

#include "../../util/stencil.hpp"
#include "../PDEModelConcepts.hpp"

namespace treesap {

template <typename RHSType, int order_ = RHSType::ngrow>
    requires RHSConcept<RHSType>
struct KODecorator {
    static constexpr int ncomp = RHSType::ncomp;
    static constexpr int ngrow = RHSType::ngrow;
    static constexpr int order = order_;
    static_assert(order > 0 && order <= ngrow, "Kreiss-Oliger order must be in 1..ngrow");

    using Workspace = typename RHSType::Workspace;
    using inner_rhs_t = RHSType;

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real time, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &workspace) const {
        amrex::Real ko = 0.0;
        ko += stencil::kreiss_oliger<order, 0>(arr, comp, idx, geom);
        ko += stencil::kreiss_oliger<order, 1>(arr, comp, idx, geom);
        ko += stencil::kreiss_oliger<order, 2>(arr, comp, idx, geom);
        return rhs.template operator()<comp>(idx, time, arr, geom, workspace) + strength * ko;
    }

    RHSType rhs;
    amrex::Real strength;
};

} // namespace treesap
