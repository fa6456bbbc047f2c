// ABLDecorator.hpp -- absorbing boundary layer: damps every component towards
// a target value inside a layer of `thickness` cells along the domain faces
// (pde-model/rhs-decorators/ABLDecorator.hpp:24-132):
//   F_c - sigma(idx) * (u_c(idx) - target_c)
//   sigma = 0 inside bulk = Domain grown by -thickness, else
//   sigma = max_strength * profile(clamp(inv_thickness * |idx - clamp(idx, bulk)|, 0, 1))
// with |.| the Euclidean norm of the integer offset (sqrt of the integer sum of squares).
#pragma once

#include "../../util/stencil.hpp"
#include "../PDEModelConcepts.hpp"

#include <cmath>
#include <utility>

namespace treesap {

// A damping profile p(s) on [0, 1] with p(0) = 0 and p(1) = 1, checked at compile time.
template <typename F>
concept ABLProfileConcept = std::is_trivially_copyable_v<F> && requires(F f, amrex::Real s) {
    { f(s) } -> std::convertible_to<amrex::Real>;
    requires F{}(0.0) == 0.0;
    requires F{}(1.0) == 1.0;
};

template <typename RHSType, typename ABLProfileType>
    requires RHSConcept<RHSType> && ABLProfileConcept<ABLProfileType>
struct ABLDecorator {
    static constexpr int ncomp = RHSType::ncomp;
    static constexpr int ngrow = RHSType::ngrow;
    using Workspace = typename RHSType::Workspace;
    using inner_rhs_t = RHSType;
    using profile_t = ABLProfileType;

    struct ABLData {
        ABLData(amrex::GpuArray<amrex::Real, ncomp> target_values_, int thickness_, amrex::Real max_strength_,
                ABLProfileType profile_)
            : profile(std::move(profile_)), thickness(thickness_),
              inv_thickness(1.0 / static_cast<amrex::Real>(thickness_)), max_strength(max_strength_),
              target_values(std::move(target_values_)) {}

        ABLProfileType profile;
        int thickness;
        amrex::Real inv_thickness;
        amrex::Real max_strength;
        amrex::GpuArray<amrex::Real, ncomp> target_values;
    };

    ABLDecorator(RHSType rhs_, const amrex::Geometry &geom, ABLData data_)
        : rhs(std::move(rhs_)), data(std::move(data_)), bulk(amrex::grow(geom.Domain(), -data.thickness)) {}

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real time, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &workspace) const {
        return rhs.template operator()<comp>(idx, time, arr, geom, workspace) -
               sigma(idx) * (arr(idx, comp) - data.target_values[comp]);
    }

    AMREX_GPU_HOST_DEVICE amrex::Real sigma(const amrex::IntVect &idx) const {
        if (bulk.contains(idx)) return 0.0;
        int d2 = 0;
        for (int a = 0; a < 3; ++a) {
            const int off = idx[a] - amrex::Clamp(idx[a], bulk.smallEnd(a), bulk.bigEnd(a));
            d2 += off * off;
        }
        const amrex::Real dist = std::sqrt(static_cast<amrex::Real>(d2));
        const amrex::Real s = amrex::Clamp(data.inv_thickness * dist, 0.0, 1.0);
        return data.max_strength * data.profile(s);
    }

    RHSType rhs;
    ABLData data;
    amrex::Box bulk;
};

} // namespace treesap
