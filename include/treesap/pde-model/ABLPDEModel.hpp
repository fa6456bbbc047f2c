// ABLPDEModel.hpp -- a PDEModel whose right-hand side is wrapped in an
// ABLDecorator and whose boundary functor is a DirichletBC holding the ABL's
// target values (pde-model/ABLPDEModel.hpp:29-51). To add dissipation as well,
// pass a KODecorator<UserRHS> as RHSType: the result is ABL(KO(user)).
#pragma once

#include "../init/ConstInit.hpp"
#include "PDEModel.hpp"
#include "bcs/DirichletBC.hpp"
#include "rhs-decorators/ABLDecorator.hpp"

namespace treesap {

template <typename RHSType, typename InitType, typename ABLProfileType>
class ABLPDEModel
    : public PDEModel<ABLDecorator<RHSType, ABLProfileType>, InitType, DirichletBC<RHSType::ncomp, RHSType::ngrow>> {
    using Decorated = ABLDecorator<RHSType, ABLProfileType>;
    using BC = DirichletBC<RHSType::ncomp, RHSType::ngrow>;
    using Base = PDEModel<Decorated, InitType, BC>;

  public:
    ABLPDEModel(amrex::Vector<std::string> comp_names, amrex::Geometry geom, amrex::Real dt, RHSType rhs,
                typename Decorated::ABLData abl_data, InitType init)
        : Base(std::move(comp_names), geom, dt, Decorated{std::move(rhs), geom, abl_data}, std::move(init),
               BC{abl_data.target_values}) {}
};

namespace detail {
struct IdentityProfile {
    constexpr amrex::Real operator()(amrex::Real s) const { return s; }
};
} // namespace detail
static_assert(PDEModelConcept<ABLPDEModel<NoOpRHS<1, 1>, ConstInit<1>, detail::IdentityProfile>>);

} // namespace treesap
