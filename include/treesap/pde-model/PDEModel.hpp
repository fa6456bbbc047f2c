// PDEModel.hpp -- the model a user hands to an integrator: right-hand side,
// initial condition, boundary functor, boundary records, geometry, component
// names and time step (pde-model/PDEModel.hpp:34-136). The default BC is NoOpBC
// and the default boundary records are interior (int_dir) on every face.
#pragma once

#include "../init/ConstInit.hpp"
#include "PDEModelConcepts.hpp"
#include "bcs/NoOpBC.hpp"
#include "rhs-decorators/NoOpRHS.hpp"

#include <string>
#include <utility>

namespace treesap {

template <typename RHSType, typename InitType, typename BCType = NoOpBC<RHSType::ncomp, RHSType::ngrow>>
    requires RHSConcept<RHSType> && BCConcept<BCType> && InitConcept<InitType>
class PDEModel {
  public:
    static constexpr int ncomp = RHSType::ncomp;
    static constexpr int ngrow = RHSType::ngrow;
    using rhs_t = RHSType;
    using init_t = InitType;
    using bc_t = BCType;

    static_assert(InitType::ncomp == ncomp, "initial condition and RHS disagree on ncomp");
    static_assert(BCType::ncomp == ncomp, "boundary functor and RHS disagree on ncomp");
    static_assert(BCType::ngrow == ngrow, "boundary functor and RHS disagree on ngrow");

    PDEModel(amrex::Vector<std::string> comp_names_, amrex::Geometry geom_, amrex::Real dt_, RHSType rhs_,
             InitType init_, BCType bc_ = BCType{},
             amrex::Vector<amrex::BCRec> bc_recs_ = amrex::Vector<amrex::BCRec>(
                 ncomp, amrex::BCRec(amrex::BCType::int_dir, amrex::BCType::int_dir, amrex::BCType::int_dir,
                                     amrex::BCType::int_dir, amrex::BCType::int_dir, amrex::BCType::int_dir)))
        : rhs(std::move(rhs_)), init(std::move(init_)), bc(std::move(bc_)), bc_recs(std::move(bc_recs_)),
          geom(std::move(geom_)), comp_names(std::move(comp_names_)), dt(dt_) {}

    const RHSType &get_rhs() const { return rhs; }
    const InitType &get_init() const { return init; }
    const BCType &get_bc() const { return bc; }
    const amrex::Vector<amrex::BCRec> &get_bc_recs() const { return bc_recs; }
    const amrex::Geometry &get_geometry() const { return geom; }
    const amrex::Vector<std::string> &get_comp_names() const { return comp_names; }
    amrex::Real dt_hint() const { return dt; }

  protected:
    RHSType rhs;
    InitType init;
    BCType bc;
    amrex::Vector<amrex::BCRec> bc_recs;
    amrex::Geometry geom;
    amrex::Vector<std::string> comp_names;
    amrex::Real dt;
};

static_assert(PDEModelConcept<PDEModel<NoOpRHS<1, 1>, ConstInit<1>>>);

} // namespace treesap
