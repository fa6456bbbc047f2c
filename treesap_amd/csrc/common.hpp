// common.hpp -- definitions shared by the HIP kernels and the C-ABI host code.
//
// Stencil weights are the closed forms of util/stencil_impl.hpp:56-140 and
// util/cexprmath.hpp:32-105, evaluated at compile time (IEEE double, round to
// nearest) so every kernel carries them as immediates. The integer factors of
// the Kreiss-Oliger operator follow util/stencil.hpp:96-103.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "treesap_hip.h"

namespace tsp {

// ---------------------------------------------------------------------------
// compile-time unrolling (util/static_for.hpp:38-68 fixes the component order)
// ---------------------------------------------------------------------------
template <int Start, int End, class F> __host__ __device__ __forceinline__ constexpr void static_for(F &&f) {
    if constexpr (Start < End) {
        f.template operator()<Start>();
        static_for<Start + 1, End>(f);
    }
}

// ---------------------------------------------------------------------------
// integer/double helpers (util/cexprmath.hpp)
// ---------------------------------------------------------------------------
namespace cx {
// cexprmath.hpp:32-37
constexpr int64_t factorial(int n) {
    int64_t res = 1;
    for (int i = 2; i <= n; ++i) res *= i;
    return res;
}
// cexprmath.hpp:46-59
constexpr int64_t binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    if (k == 0 || k == n) return 1;
    if (k > n / 2) k = n - k;
    int64_t res = 1;
    for (int i = 1; i <= k; ++i) res = (res * (n - i + 1)) / i;
    return res;
}
// cexprmath.hpp:68-77
constexpr int64_t power(int64_t n, int k) {
    int64_t result = 1;
    while (k > 0) {
        if (k & 1) result *= n;
        n *= n;
        k >>= 1;
    }
    return result;
}
// cexprmath.hpp:85
constexpr int alternating_sign(int k) { return (k & 1) ? -1 : 1; }
// cexprmath.hpp:99-105
constexpr double harmonic(int n, int m) {
    double sum = 0.0;
    for (int k = 1; k <= n; k++) sum += 1.0 / static_cast<double>(power(k, m));
    return sum;
}
// stencil_impl.hpp:56-71
constexpr double w1(int n, int p) {
    const int ka = p > 0 ? p : -p;
    if (p == 0) return 0;
    const int64_t num = alternating_sign(ka + 1) * power(factorial(n), 2);
    const int64_t den = ka * factorial(n - ka) * factorial(n + ka);
    const double w = static_cast<double>(num) / static_cast<double>(den);
    return p > 0 ? w : -w;
}
// stencil_impl.hpp:91-104
constexpr double w2(int n, int p) {
    const int k = p > 0 ? p : -p;
    if (p == 0) return -2.0 * harmonic(n, 2);
    const int64_t num = 2 * alternating_sign(k + 1) * power(factorial(n), 2);
    const int64_t den = k * k * factorial(n - k) * factorial(n + k);
    return static_cast<double>(num) / static_cast<double>(den);
}
// stencil_impl.hpp:124-140
constexpr double w3(int n, int p) {
    const int ka = p > 0 ? p : -p;
    if (p == 0) return 0;
    const int64_t num = 6 * alternating_sign(ka + 1) * power(factorial(n), 2);
    const int64_t den = power(ka, 3) * factorial(n - ka) * factorial(n + ka);
    const double w = ((1.0 - static_cast<double>(ka * ka) * harmonic(n, 2)) * static_cast<double>(num) /
                      static_cast<double>(den));
    return p > 0 ? w : -w;
}
// stencil.hpp:96-101: centre -binom(2r,r); offset s: -alternating_sign(s)*binom(2r,r+s)
constexpr int64_t ko_coef(int r, int s) {
    if (s == 0) return -binom(2 * r, r);
    return -alternating_sign(s) * binom(2 * r, r + s);
}
// stencil.hpp:93
constexpr double ko_norm(int r) { return 1.0 / static_cast<double>(power(2, 2 * r)); }
} // namespace cx

// ---------------------------------------------------------------------------
// stage arguments (one RK4 stage or one Euler update over one box)
// ---------------------------------------------------------------------------
enum StageKind : int32_t { kStage0 = 0, kStage1 = 1, kStage2 = 2, kStage3 = 3, kEuler = 4 };

struct StageArgs {
    const double *src;       // register read with stencils (init for stage 0, arg otherwise)
    const double *init;      // state_old (centre reads)
    double *acc;             // state_new
    double *argn;            // arg_next
    double *kko;             // tensor system with KO: KO sums of the K components (register layout, 6 components)
    const double *sigma_tab; // ABL sigma by squared distance to the bulk box
    int32_t nx, ny, nz;
    int32_t G;               // ghost width of the allocation
    int32_t sj, sk, sc;      // element strides of the register layout
    int32_t stage;           // StageKind
    int32_t ghost_mode;      // 0 virtual ghosts, 1 read ghost memory
    int32_t periodic[3];
    int32_t dirichlet;
    int32_t abl_T;           // ABL thickness (0: no ABL decorator)
    int32_t off[3];          // global index of the box's cell (0,0,0) within the domain (a box of a level)
    int32_t dom[3];          // cells per axis of the domain (the absorbing layer is measured against it)
    int32_t zchunk;
    int32_t tiles_x, tiles_y, nzc;
    int32_t xcd_remap;
    int32_t zalt;            // alternate the z-march direction of consecutive chunks
    int32_t halo_ring;       // 64x8/64x16 tiles: load halos with the queue head, park them in LDS
    int32_t ptr_access;      // 1: 64-bit pointer addressing (registers of 4 GiB or more)
    double b, a_next;        // dt*b_s and dt*a_{s+1} (integrators/SingleLevelRK4Integrator.hpp:119-121)
    double c_init_argn;      // adjoint only: coefficient of lambda' in the next stage argument
    double src_scale;        // adjoint only: factor applied to every source value on load
    double b1, a_next1;      // stage-pair kernel: dt*b and dt*a_next of the pair's second stage
    double gf1[3];           // 1/dx   (stencil_impl.hpp:213, stencil.hpp:94)
    double gf2[3];           // 1/(dx*dx) (stencil_impl.hpp:255)
    double ko_strength;
    double gamma;
    double target[TSP_MAX_COMP];
};

// Launch a fused stage kernel of one (system, stencil width, KO order) with
// tile height `tile_y`; fills the tile counts of a copy of `a`.
using LaunchFn = hipError_t (*)(const StageArgs &a, int tile_y, hipStream_t stream);

LaunchFn lookup_phipi(int n, int r);
LaunchFn lookup_phipi_adjoint(int n, int r);
LaunchFn lookup_phiu(int n, int r);
LaunchFn lookup_nonlinear(int n, int r);
LaunchFn lookup_tensor(int n, int r);
// Launch a stage-pair kernel (forward RK4 stages first, first+1 in one launch;
// first = 0 or 2); hipErrorNotSupported, without launching, where it does not apply.
using PairFn = hipError_t (*)(const StageArgs &a, int first, hipStream_t stream);
PairFn lookup_pair_phipi(int n, int r);
// static LDS bytes of the (phi, Pi) stage-pair kernel (0 where none is built)
size_t pair_lds_bytes_phipi(int n, int r);
// launchers specialised for forward RK4 stage 0 / stage 3 / the middle stages
// 1 and 2 (hipErrorNotSupported for configurations they do not cover: the
// caller then uses the generic one)
LaunchFn lookup_phipi_s0(int n, int r);
LaunchFn lookup_phipi_s1(int n, int r);
LaunchFn lookup_phipi_s3(int n, int r);
LaunchFn lookup_phiu_s0(int n, int r);
LaunchFn lookup_phiu_s3(int n, int r);
LaunchFn lookup_nonlinear_s0(int n, int r);
LaunchFn lookup_nonlinear_s3(int n, int r);

// Ghost fill (SingleLevelIntegrator.hpp:153-163) of one register.
struct GhostArgs {
    double *p; // element (0,0,0,0)
    int32_t nx, ny, nz, G, nc;
    int32_t sj, sk, sc;
    int32_t periodic[3];
    int32_t dirichlet;
    double target[TSP_MAX_COMP];
};
hipError_t launch_ghost_fill(const GhostArgs &g, hipStream_t stream);

// nonlinear terms of one reverse stage of the NonLinear system's discrete
// adjoint (adjoint_nl.hip): acc += N^T mu, argn += a_next N^T mu, N the
// linearised -gamma pi |grad phi|^2 at the stage argument abar
struct NlAdjArgs {
    const double *mu;   // mu_j, register origin (element (0,0,0,0))
    double mu_scale;    // mu_j = mu_scale * register (b_3 at the last stage, else 1)
    const double *abar; // the forward stage argument a_j (phi, pi), ghosts as the forward stage read them
    double *acc;        // sum of the stage terms
    double *argn;       // the next stage's mu (null at stage 0)
    double a_next;      // a_j
    double *w;          // 4 dense scratch fields over the valid cells
    int32_t nx, ny, nz, sj, sk, sc, n;
    int32_t periodic[3];
    double gf1[3];
    double gamma;
};
hipError_t launch_nonlinear_adjoint_terms(const NlAdjArgs &a, hipStream_t stream);

// valid-region pack/unpack between a dense [c][k][j][i] buffer and a register
hipError_t launch_unpack(double *reg0, const double *dense, int nx, int ny, int nz, int nc, int sj, int sk, int sc,
                         hipStream_t stream);
hipError_t launch_pack(double *dense, const double *reg0, int nx, int ny, int nz, int nc, int sj, int sk, int sc,
                       hipStream_t stream);

// receiver traces of the adjoint solver
// (rows outside [0, nrows) are skipped)
hipError_t launch_gather_trace(const double *phi, const int *rec_off, int nrec, double *trace, int nrows,
                               int *counter, hipStream_t stream);
hipError_t launch_inject_trace(double *lam, const int *rec_off, int nrec, const double *dtrace, int nrows,
                               int *counter, hipStream_t stream);
hipError_t launch_set_counter(int *counter, int v, hipStream_t stream);
hipError_t launch_zero(double *p, long long n, hipStream_t stream);

// dynamic LDS bytes for the ABL rate table of thickness T
inline size_t sigma_lds_bytes(int T) { return T > 0 ? sizeof(double) * (size_t)(3 * T * T + 1) : 0; }

// LDS a block may hold (static + dynamic), queried once per device
inline size_t max_block_lds() {
    static thread_local int dev_cached = -1;
    static thread_local size_t bytes = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (dev != dev_cached) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 0;
        bytes = (size_t)v;
        dev_cached = dev;
    }
    return bytes;
}


} // namespace tsp
