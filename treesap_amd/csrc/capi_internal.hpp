// capi_internal.hpp -- host-side helpers shared by the single-box (capi.hip)
// and batched (batch.hip) halves of the C ABI: error reporting, model
// validation, register layout and the tabulated ABL damping rate.
#pragma once

#include "common.hpp"

#include <cmath>
#include <string>
#include <vector>

namespace tsp_host {

// thread-local message behind tsp_last_error()
extern thread_local std::string g_err;

inline tsp_status fail(tsp_status st, const std::string &msg) {
    g_err = msg;
    return st;
}

inline tsp_status hip_fail(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return TSP_HIP_ERROR;
}

#define TSP_HIP(call)                                                                                                  \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return ::tsp_host::hip_fail(e_, #call);                                                  \
    } while (0)

constexpr int kXOff = 16; // valid x starts 16 doubles (128 B) into each row

// integrators/SingleLevelRK4Integrator.hpp:71-74
constexpr double kAFrac[4] = {0.0, 0.5, 0.5, 1.0};
constexpr double kBFrac[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

inline int system_ncomp(int system) {
    switch (system) {
    case TSP_SYS_WAVE_PHIPI: return 2;
    case TSP_SYS_WAVE_PHIU: return 4;
    case TSP_SYS_NONLINEAR: return 2;
    case TSP_SYS_TENSOR_ADM: return 12;
    default: return -1;
    }
}

// component count of a model (a custom model states its own)
inline int desc_ncomp(const tsp_desc &d) {
    if (d.system == TSP_SYS_CUSTOM) return d.ncomp >= 1 && d.ncomp <= TSP_MAX_COMP ? d.ncomp : -1;
    return system_ncomp(d.system);
}

inline tsp::LaunchFn lookup_launcher(int system, int n, int r) {
    switch (system) {
    case TSP_SYS_WAVE_PHIPI: return tsp::lookup_phipi(n, r);
    case TSP_SYS_WAVE_PHIU: return tsp::lookup_phiu(n, r);
    case TSP_SYS_NONLINEAR: return tsp::lookup_nonlinear(n, r);
    case TSP_SYS_TENSOR_ADM: return tsp::lookup_tensor(n, r);
    default: return nullptr;
    }
}

// Launcher specialised for forward RK4 stage `stage`, or null.
inline tsp::LaunchFn lookup_stage_launcher(int system, int n, int r, int stage) {
    switch (system) {
    case TSP_SYS_WAVE_PHIPI:
        return stage == 0 ? tsp::lookup_phipi_s0(n, r) : stage == 3 ? tsp::lookup_phipi_s3(n, r) : tsp::lookup_phipi_s1(n, r);
    case TSP_SYS_WAVE_PHIU: return stage == 0 ? tsp::lookup_phiu_s0(n, r) : stage == 3 ? tsp::lookup_phiu_s3(n, r) : nullptr;
    case TSP_SYS_NONLINEAR:
        return stage == 0 ? tsp::lookup_nonlinear_s0(n, r) : stage == 3 ? tsp::lookup_nonlinear_s3(n, r) : nullptr;
    default: return nullptr;
    }
}

// ABLProfileConcept instances (rhs-decorators/ABLDecorator.hpp:24-29)
inline double abl_profile(int id, double s) {
    switch (id) {
    case TSP_PROFILE_LINEAR: return s;
    case TSP_PROFILE_QUADRATIC: return s * s;
    default: return s * s * s;
    }
}

// ABLDecorator.hpp:115-132 tabulated by squared integer distance d2 to the
// bulk box (d2 = 0 is the bulk, sigma 0); 3*T^2 + 1 entries.
inline std::vector<double> abl_sigma_table(const tsp_desc &d) {
    const int T = d.abl_thickness;
    const int nt = 3 * T * T + 1;
    std::vector<double> tab(nt, 0.0);
    const double inv_thickness = 1.0 / static_cast<double>(T);
    for (int d2 = 1; d2 < nt; ++d2) {
        const double norm = std::sqrt(static_cast<double>(d2));
        const double x = inv_thickness * norm;
        const double scaled = (x < 0.0) ? 0.0 : (1.0 < x) ? 1.0 : x;
        tab[d2] = d.abl_max * abl_profile(d.abl_profile, scaled);
    }
    return tab;
}

// The stage kernels keep the ABL rate table in LDS beside their plane buffers
// (at most 56 KiB: 64 x 16 tiles, two staged components, halo 4, double
// buffered); reject layers whose table would not fit the device's LDS.
inline tsp_status check_abl_lds(int T) {
    if (T <= 0) return TSP_OK;
    int dev = 0, max_lds = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute(MaxSharedMemoryPerBlock)");
    const long long need = 56LL * 1024 + (long long)sizeof(double) * (3LL * T * T + 1);
    if (need > max_lds)
        return fail(TSP_UNSUPPORTED, "ABL thickness " + std::to_string(T) + " needs a " +
                                         std::to_string(sizeof(double) * (3LL * T * T + 1)) +
                                         "-byte rate table in LDS beside the 56 KiB plane buffers; the device has " +
                                         std::to_string(max_lds) + " bytes per workgroup");
    return TSP_OK;
}

// Register layout of one box: pitch P (multiple of 16 doubles), strides in elements.
struct Layout {
    int G = 0, pitch = 0;
    int64_t sj = 0, sk = 0, sc = 0, elems = 0; // elems: whole register, all components
};
inline Layout layout_of(const tsp_desc &d) {
    Layout L;
    L.G = d.ghost;
    L.pitch = (int)(((int64_t)kXOff + d.n[0] + L.G + 15) / 16 * 16);
    L.sj = L.pitch;
    L.sk = L.sj * (d.n[1] + 2 * L.G);
    L.sc = L.sk * (d.n[2] + 2 * L.G);
    L.elems = L.sc * desc_ncomp(d);
    return L;
}

inline tsp_status validate(const tsp_desc *d) {
    if (!d) return fail(TSP_INVALID_ARG, "descriptor is NULL");
    if (d->system == TSP_SYS_CUSTOM && desc_ncomp(*d) < 0)
        return fail(TSP_INVALID_ARG, "a custom model needs ncomp in 1..12");
    const int nc = desc_ncomp(*d);
    if (nc < 0) return fail(TSP_INVALID_ARG, "unknown system id " + std::to_string(d->system));
    for (int a = 0; a < 3; ++a)
        if (d->n[a] < 1) return fail(TSP_INVALID_ARG, "n[" + std::to_string(a) + "] must be >= 1");
    const bool custom = d->system == TSP_SYS_CUSTOM;
    if (custom && (d->ko_order != 0 || d->abl_thickness != 0))
        return fail(TSP_INVALID_ARG, "a custom model's KO and ABL decorators are part of its functor "
                                     "(ko_order and abl_thickness must be 0)");
    if (!custom && (d->stencil_n < 1 || d->stencil_n > 4)) return fail(TSP_INVALID_ARG, "stencil_n must be in 1..4");
    if (d->ko_order < 0 || d->ko_order > 4) return fail(TSP_INVALID_ARG, "ko_order must be in 0..4");
    if (d->ghost < 0 || (!custom && (d->ghost < d->stencil_n || d->ghost < d->ko_order)))
        return fail(TSP_INVALID_ARG, "ghost width must be >= stencil_n and >= ko_order (KODecorator.hpp:43)");
    if (d->ghost > kXOff) return fail(TSP_INVALID_ARG, "ghost width must be <= 16");
    if (!custom && !lookup_launcher(d->system, d->stencil_n, d->ko_order))
        return fail(TSP_UNSUPPORTED, "no stage kernel for stencil_n=" + std::to_string(d->stencil_n) +
                                         " ko_order=" + std::to_string(d->ko_order) +
                                         (d->system == TSP_SYS_TENSOR_ADM ? " (supported: ko_order in {0, n})"
                                                                          : " (supported: ko_order in {0, n, n+1}, <= 4)"));
    for (int a = 0; a < 3; ++a) {
        if (!(d->prob_hi[a] > d->prob_lo[a])) return fail(TSP_INVALID_ARG, "prob_hi must exceed prob_lo");
        if (d->periodic[a] != 0 && d->periodic[a] != 1) return fail(TSP_INVALID_ARG, "periodic flags must be 0 or 1");
        if (d->periodic[a] && d->ghost > d->n[a])
            return fail(TSP_INVALID_ARG, "periodic axis needs n >= ghost width");
    }
    if (d->bc != TSP_BC_NOOP && d->bc != TSP_BC_DIRICHLET && d->bc != TSP_BC_CUSTOM)
        return fail(TSP_INVALID_ARG, "unknown bc id");
    if (d->abl_thickness < 0) return fail(TSP_INVALID_ARG, "abl_thickness must be >= 0");
    if (d->abl_thickness > 0) {
        for (int a = 0; a < 3; ++a)
            if (2 * d->abl_thickness >= d->n[a])
                return fail(TSP_INVALID_ARG, "ABL layer leaves an empty bulk box (2*thickness >= n)");
        if (d->abl_profile < TSP_PROFILE_LINEAR || d->abl_profile > TSP_PROFILE_CUBIC)
            return fail(TSP_INVALID_ARG, "unknown ABL profile id");
        if (!std::isfinite(d->abl_max)) return fail(TSP_INVALID_ARG, "abl_max must be finite");
    }
    if (!std::isfinite(d->dt)) return fail(TSP_INVALID_ARG, "dt must be finite");
    if (!std::isfinite(d->ko_strength)) return fail(TSP_INVALID_ARG, "ko_strength must be finite");
    const int G = d->ghost;
    const int64_t pitch = ((int64_t)kXOff + d->n[0] + G + 15) / 16 * 16;
    const int64_t elems = pitch * (d->n[1] + 2 * G) * (int64_t)(d->n[2] + 2 * G) * nc;
    if (elems >= (1LL << 31)) return fail(TSP_UNSUPPORTED, "register exceeds 2^31 elements");
    return TSP_OK;
}

// initial data from postfix expression programs (init_expr.hip)
tsp_status check_expr_programs(const tsp_expr_op *ops, const int32_t *comp_start, int nc);

} // namespace tsp_host

namespace tsp {
// one box whose state_old valid cells an expression program fills
struct InitBox {
    double *origin; // element (0, 0, 0, 0) of the box's state_old register
    long long sj, sk, sc;
    int n[3];
    int off[3]; // the box's first cell, domain index relative to the domain's first cell
};
} // namespace tsp

namespace tsp_host {
// evaluate the (checked) programs into the boxes on `stream` and wait; `domain`
// gives prob_lo, prob_hi and the domain's cell counts
tsp_status run_init_expr(const std::vector<tsp::InitBox> &boxes, const tsp_desc &domain, const tsp_expr_op *ops,
                         const int32_t *comp_start, int nc, const int32_t index_lo[3], hipStream_t stream);
} // namespace tsp_host
