// treesap.hpp -- the whole C++ drop-in API in one include: TreESaP's plug-in
// types (PDEModel, ABLPDEModel, concepts, decorators, boundary functors,
// initial conditions, tensor aliases and operations, stencils) and the
// single-level RK4 and Euler integrators, which run on the MI355X through the
// C ABI of include/treesap_hip.h.
//
// Build: g++ -std=c++20 -I<repo>/include/treesap -I<repo>/include app.cpp
//        -L<repo>/treesap_amd -ltreesap_hip -Wl,-rpath,<repo>/treesap_amd
// (the include paths of the reference, e.g. "pde-model/PDEModel.hpp", resolve
// against include/treesap).
#pragma once

#include "amrex_types.hpp"

#include "Integrator.hpp"
#include "SingleLevelIntegrator.hpp"
#include "init/ConstInit.hpp"
#include "init/FunctionInit.hpp"
#include "integrators/SingleLevelEulerIntegrator.hpp"
#include "integrators/SingleLevelRK4Integrator.hpp"
#include "pde-model/ABLPDEModel.hpp"
#include "pde-model/PDEModel.hpp"
#include "pde-model/PDEModelConcepts.hpp"
#include "pde-model/bcs/DirichletBC.hpp"
#include "pde-model/bcs/NoOpBC.hpp"
#include "pde-model/rhs-decorators/ABLDecorator.hpp"
#include "pde-model/rhs-decorators/KODecorator.hpp"
#include "pde-model/rhs-decorators/NoOpRHS.hpp"
#include "pde-model/tensor-alias/IndexArray.hpp"
#include "pde-model/tensor-alias/RankNAlias.hpp"
#include "pde-model/tensor-alias/SymRank2Alias.hpp"
#include "pde-model/tensor-alias/TensorAliasBase.hpp"
#include "systems/registered_rhs.hpp"
#include "util/PlotExporter.hpp"
#include "util/ScopedTimer.hpp"
#include "util/cexprmath.hpp"
#include "util/named_tensors.hpp"
#include "util/static_for.hpp"
#include "util/stencil.hpp"
#include "util/tensor_operations.hpp"
#include "util/tensor_stencil.hpp"
