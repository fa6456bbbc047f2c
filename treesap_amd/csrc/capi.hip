// capi.hip -- the extern "C" boundary (include/treesap_hip.h) and solver state.
//
// A solver is one PDE model on one box: the four RK4 registers of
// integrators/SingleLevelRK4Integrator.hpp:162-166 (state_old, state_new, arg,
// arg_next) in device memory, the register-role bookkeeping of the reference's
// swaps (SingleLevelRK4Integrator.hpp:123-126, SingleLevelIntegrator.hpp:125),
// the simulation clock of Integrator.hpp:91-106, and the chosen stage-kernel
// configuration.
//
// Register layout (FP64, one allocation per register): x rows are padded so
// that the first valid cell of every row sits on a 128-byte boundary:
//   row  = [pad | G ghosts | nx valid | G ghosts | pad],  pitch P (multiple of 16),
//   element (i,j,k,c) at base + (XOFF+i) + (G+j)*P + (G+k)*P*(ny+2G) + c*P*(ny+2G)*(nz+2G).
#include "capi_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

namespace tsp_host {
thread_local std::string g_err;
} // namespace tsp_host

using namespace tsp_host;

namespace {
} // namespace

struct tsp_solver {
    tsp_desc d{};
    int nc = 0, G = 0, H = 0;
    int pitch = 0;
    int64_t sj = 0, sk = 0, sc = 0, reg_elems = 0;
    double *base[4] = {nullptr, nullptr, nullptr, nullptr};
    int role[4] = {0, 1, 2, 3}; // role -> register index
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = -1; // the device current at creation
    // TSP_POINTER_ACCESS set at creation: run the pointer-access stage kernel
    // that registers of 4 GiB or more use, on any size (for testing it)
    bool ptr_access = false;
    double *sigma_tab = nullptr;
    // tensor system with KO: the K components' KO sums of the current stage
    // (six components in the register layout, written and read by the stage launch)
    double *kko = nullptr;
    // TSP_BC_CUSTOM: the caller's physical-boundary callback, run after each periodic fill
    tsp_bc_fn bc_fn = nullptr;
    void *bc_user = nullptr;
    tsp::LaunchFn launch = nullptr;
    tsp::LaunchFn launch_stage[4] = {nullptr, nullptr, nullptr, nullptr}; // stage-specific, where built
    tsp::PairFn launch_pair = nullptr;                                      // stage pairs, where built
    tsp_kernel_cfg cfg{0, 8, 64, 1, 1, 1, 0};
    double time = 0.0;
    int64_t steps = 0;
    double gf1[3]{}, gf2[3]{};
    // hipGraphs of two RK4 steps, each for the register roles it was captured
    // with (two steps bring the roles back), valid for `graph_cfg`
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_role[2][4] = {};
    int graph_next = 0; // slot the next capture replaces when both are taken
    tsp_kernel_cfg graph_cfg{};
};

namespace {

// Default stage-kernel configuration for a box: tall tiles where the box is
// tall enough, and the longest z chunk (of 64, 32, 16, 8) that still gives at
// least two workgroups per CU, so small boxes fill the device too.
tsp_kernel_cfg default_cfg(const tsp_desc &d) {
    // the tensor kernel: 16 x 8 x 2 blocks, block order in slabs of two z tiles
    if (d.system == TSP_SYS_TENSOR_ADM) return tsp_kernel_cfg{1, 8, 2, 1, 0, 0, 0};
    tsp_kernel_cfg c{0, d.n[1] >= 16 ? 16 : 8, 64, 1, 1, 1, 0};
    // mixed stencils read filled edge/corner ghosts; a custom kernel reads ghost
    // memory, and so does every kernel under a custom boundary functor
    if (d.system == TSP_SYS_TENSOR_ADM || d.system == TSP_SYS_CUSTOM || d.bc == TSP_BC_CUSTOM) c.ghost_mode = 1;
    const long long tiles = (long long)((d.n[0] + 63) / 64) * ((d.n[1] + c.tile_y - 1) / c.tile_y);
    for (int zc : {64, 32, 16, 8}) {
        c.zchunk = std::min(zc, d.n[2]);
        if (tiles * ((d.n[2] + c.zchunk - 1) / c.zchunk) >= 512) break;
    }
    return c;
}

// Whether a step runs the fill pass before its stages. Under a custom
// boundary functor it always does. The tensor system's stage kernel and a
// custom stage kernel read ghost memory: they need the pass for periodic
// images; without a periodic axis every ghost cell lies outside a
// non-periodic face and holds what the pass would write there (the Dirichlet
// values written at creation, or, with NoOp, memory no stage writes: stages
// write valid cells only). The other kernels need it in ghost_mode 1.
bool fills_ghosts(const tsp_solver *s) {
    if (s->d.bc == TSP_BC_CUSTOM) return true;
    if (s->d.system == TSP_SYS_TENSOR_ADM || s->d.system == TSP_SYS_CUSTOM)
        return s->d.periodic[0] || s->d.periodic[1] || s->d.periodic[2];
    return s->cfg.ghost_mode == 1;
}

double *reg_origin(const tsp_solver *s, int reg) {
    return s->base[reg] + kXOff + (int64_t)s->G * s->sj + (int64_t)s->G * s->sk;
}

tsp_status validate_cfg(const tsp_kernel_cfg *c) {
    if (!c) return fail(TSP_INVALID_ARG, "kernel config is NULL");
    if (c->ghost_mode != 0 && c->ghost_mode != 1) return fail(TSP_INVALID_ARG, "ghost_mode must be 0 or 1");
    if (c->tile_y != 2 && c->tile_y != 4 && c->tile_y != 8 && c->tile_y != 16 && c->tile_y != 32)
        return fail(TSP_INVALID_ARG, "tile_y must be 2, 4, 8, 16 or 32");
    if (c->zchunk < 1) return fail(TSP_INVALID_ARG, "zchunk must be >= 1");
    if (c->xcd_remap != 0 && c->xcd_remap != 1) return fail(TSP_INVALID_ARG, "xcd_remap must be 0 or 1");
    if (c->zalt < 0 || c->zalt > 2) return fail(TSP_INVALID_ARG, "zalt must be 0, 1 or 2");
    if (c->halo_ring != 0 && c->halo_ring != 1) return fail(TSP_INVALID_ARG, "halo_ring must be 0 or 1");
    if (c->stage_pairs != 0 && c->stage_pairs != 1) return fail(TSP_INVALID_ARG, "stage_pairs must be 0 or 1");
    return TSP_OK;
}

// Whether the stage-pair kernels apply to this solver with ghost mode
// `ghost_mode` (the launcher re-checks what depends on the device).
bool pairs_apply(const tsp_solver *s, int ghost_mode, std::string *why) {
    auto no = [&](const char *m) {
        if (why) *why = m;
        return false;
    };
    if (!s->launch_pair) return no("stage_pairs needs the (phi, Pi) system with KO dissipation");
    if (ghost_mode != 0) return no("stage_pairs needs ghost_mode 0");
    if (s->d.bc == TSP_BC_CUSTOM) return no("stage_pairs needs a boundary functor the library carries");
    if (s->ptr_access || (int64_t)sizeof(double) * s->reg_elems >= (int64_t(1) << 32))
        return no("stage_pairs needs registers under 4 GiB");
    for (int a = 0; a < 3; ++a)
        if (s->d.periodic[a] && s->d.n[a] < 2 * s->H) return no("stage_pairs needs periodic axes of at least 2H cells");
    if (tsp::pair_lds_bytes_phipi(s->d.stencil_n, s->d.ko_order) + tsp::sigma_lds_bytes(s->d.abl_thickness) >
        tsp::max_block_lds())
        return no("stage_pairs: the pair kernel's LDS and the ABL rate table exceed the device's LDS per workgroup");
    return true;
}

void drop_graphs(tsp_solver *s) {
    for (int p = 0; p < 2; ++p)
        if (s->graph[p]) {
            (void)hipGraphExecDestroy(s->graph[p]);
            s->graph[p] = nullptr;
        }
}

tsp::GhostArgs ghost_args(const tsp_solver *s, int reg) {
    tsp::GhostArgs g{};
    g.p = reg_origin(s, reg);
    g.nx = s->d.n[0];
    g.ny = s->d.n[1];
    g.nz = s->d.n[2];
    g.G = s->G;
    g.nc = s->nc;
    g.sj = (int32_t)s->sj;
    g.sk = (int32_t)s->sk;
    g.sc = (int32_t)s->sc;
    for (int a = 0; a < 3; ++a) g.periodic[a] = s->d.periodic[a];
    g.dirichlet = s->d.bc == TSP_BC_DIRICHLET;
    for (int c = 0; c < TSP_MAX_COMP; ++c) g.target[c] = s->d.target[c];
    return g;
}

tsp::StageArgs stage_args(const tsp_solver *s, int stage, int src_reg, int init_reg, int acc_reg, int argn_reg,
                          double b, double a_next) {
    tsp::StageArgs a{};
    a.src = reg_origin(s, src_reg);
    a.init = reg_origin(s, init_reg);
    a.acc = reg_origin(s, acc_reg);
    a.argn = argn_reg >= 0 ? reg_origin(s, argn_reg) : nullptr;
    a.kko = s->kko ? s->kko + kXOff + (int64_t)s->G * s->sj + (int64_t)s->G * s->sk : nullptr;
    a.sigma_tab = s->sigma_tab;
    a.nx = s->d.n[0];
    a.ny = s->d.n[1];
    a.nz = s->d.n[2];
    a.G = s->G;
    a.sj = (int32_t)s->sj;
    a.sk = (int32_t)s->sk;
    a.sc = (int32_t)s->sc;
    a.stage = stage;
    a.ghost_mode = s->cfg.ghost_mode;
    for (int i = 0; i < 3; ++i) a.periodic[i] = s->d.periodic[i];
    a.dirichlet = s->d.bc == TSP_BC_DIRICHLET;
    a.abl_T = s->d.abl_thickness;
    for (int i = 0; i < 3; ++i) {
        a.off[i] = 0;
        a.dom[i] = s->d.n[i];
    }
    a.zchunk = s->cfg.zchunk;
    a.xcd_remap = s->cfg.xcd_remap;
    a.zalt = s->cfg.zalt;
    a.halo_ring = s->cfg.halo_ring;
    a.ptr_access = s->ptr_access ? 1 : 0;
    a.b = b;
    a.a_next = a_next;
    for (int i = 0; i < 3; ++i) {
        a.gf1[i] = s->gf1[i];
        a.gf2[i] = s->gf2[i];
    }
    a.ko_strength = s->d.ko_strength;
    a.gamma = s->d.gamma;
    for (int c = 0; c < TSP_MAX_COMP; ++c) a.target[c] = s->d.target[c];
    return a;
}

// FillPatch (SingleLevelIntegrator.hpp:153-163) of register `reg` at time
// `time`: the library's periodic (and Dirichlet) fill, then, under a custom
// boundary functor, the caller's callback on the physical-boundary ghosts.
tsp_status fill_patch(tsp_solver *s, int reg, double time) {
    TSP_HIP(tsp::launch_ghost_fill(ghost_args(s, reg), s->stream));
    if (s->d.bc != TSP_BC_CUSTOM) return TSP_OK;
    if (!s->bc_fn) return fail(TSP_INVALID_ARG, "a TSP_BC_CUSTOM solver steps only after tsp_set_bc_callback");
    tsp_bc_view v{};
    for (int a = 0; a < 3; ++a) {
        v.n[a] = s->d.n[a];
        v.periodic[a] = s->d.periodic[a];
    }
    v.ghost = s->G;
    v.ncomp = s->nc;
    v.time = time;
    v.reg = reg_origin(s, reg);
    v.sj = s->sj;
    v.sk = s->sk;
    v.sc = s->sc;
    const int r = s->bc_fn(s->bc_user, &v, (void *)s->stream);
    if (r != 0) return fail(TSP_HIP_ERROR, "boundary callback failed (returned " + std::to_string(r) + ")");
    return TSP_OK;
}

struct EventPair {
    hipEvent_t start = nullptr, stop = nullptr;
};

// Events recorded around every launch of one RK4 step: the four stage kernels
// and, in ghost_mode 1, the four fills (pre_step and before stages 1-3).
struct StepTiming {
    EventPair stage[4];
    EventPair fill[4];
};

// Enqueue one RK4 step on the solver's stream (SingleLevelIntegrator.hpp:102-125,
// SingleLevelRK4Integrator.hpp:95-157).
tsp_status enqueue_rk4_step(tsp_solver *s, StepTiming *tm) {
    if (!s->launch) return fail(TSP_UNSUPPORTED, "a custom model steps through tsp_rk4_steps_custom");
    const double dt = s->d.dt;
    int &OLD = s->role[TSP_ROLE_OLD], &NEW = s->role[TSP_ROLE_NEW], &ARG = s->role[TSP_ROLE_ARG],
        &ARGN = s->role[TSP_ROLE_ARG_NEXT];
    // slot 0: pre_step at `time`; slot s: before stage s at time + dt * a_fraction[s]
    auto fill = [&](int reg, int slot) -> tsp_status {
        if (!fills_ghosts(s)) return TSP_OK;
        if (tm) TSP_HIP(hipEventRecord(tm->fill[slot].start, s->stream));
        if (tsp_status f = fill_patch(s, reg, s->time + dt * kAFrac[slot]); f != TSP_OK) return f;
        if (tm) TSP_HIP(hipEventRecord(tm->fill[slot].stop, s->stream));
        return TSP_OK;
    };
    tsp_status st = fill(OLD, 0); // pre_step
    if (st != TSP_OK) return st;
    if (s->cfg.stage_pairs && s->launch_pair) {
        // stages (0,1) and (2,3): the registers the four single stages would
        // use, with the intermediate argument registers read at ghosts only
        const int ra = ARG, rb = ARGN;
        tsp::StageArgs p0 = stage_args(s, 0, OLD, rb, NEW, ra, dt * kBFrac[0], dt * kAFrac[1]);
        p0.b1 = dt * kBFrac[1];
        p0.a_next1 = dt * kAFrac[2];
        tsp::StageArgs p2 = stage_args(s, 2, ra, OLD, NEW, rb, dt * kBFrac[2], dt * kAFrac[3]);
        p2.b1 = dt * kBFrac[3];
        p2.a_next1 = 0.0;
        if (tm) TSP_HIP(hipEventRecord(tm->stage[0].start, s->stream));
        // tsp_set_kernel admitted the mode (pairs_apply), so the pair kernel must run
        hipError_t e = s->launch_pair(p0, 0, s->stream);
        if (e == hipErrorNotSupported) return fail(TSP_UNSUPPORTED, "stage-pair kernel does not apply to this solver");
        {
            TSP_HIP(e);
            if (tm) {
                TSP_HIP(hipEventRecord(tm->stage[0].stop, s->stream));
                TSP_HIP(hipEventRecord(tm->stage[1].start, s->stream));
                TSP_HIP(hipEventRecord(tm->stage[1].stop, s->stream));
                TSP_HIP(hipEventRecord(tm->stage[2].start, s->stream));
            }
            TSP_HIP(s->launch_pair(p2, 2, s->stream));
            if (tm) {
                TSP_HIP(hipEventRecord(tm->stage[2].stop, s->stream));
                TSP_HIP(hipEventRecord(tm->stage[3].start, s->stream));
                TSP_HIP(hipEventRecord(tm->stage[3].stop, s->stream));
            }
            std::swap(ARG, ARGN); // three swaps of the single-stage sequence
            std::swap(OLD, NEW);
            return TSP_OK;
        }
    }
    for (int stage = 0; stage < 4; ++stage) {
        const double a_next = dt * kAFrac[(stage + 1) % 4];
        const double b = dt * kBFrac[stage];
        if (stage > 0) {
            std::swap(ARG, ARGN);
            st = fill(ARG, stage);
            if (st != TSP_OK) return st;
        }
        const int src = stage == 0 ? OLD : ARG;
        const tsp::StageArgs a = stage_args(s, stage, src, OLD, NEW, stage < 3 ? ARGN : -1, b, a_next);
        if (tm) TSP_HIP(hipEventRecord(tm->stage[stage].start, s->stream));
        hipError_t e = hipErrorNotSupported;
        if (s->launch_stage[stage]) e = s->launch_stage[stage](a, s->cfg.tile_y, s->stream);
        if (e == hipErrorNotSupported) e = s->launch(a, s->cfg.tile_y, s->stream);
        TSP_HIP(e);
        if (tm) TSP_HIP(hipEventRecord(tm->stage[stage].stop, s->stream));
    }
    std::swap(OLD, NEW); // post_step
    return TSP_OK;
}

tsp_status enqueue_euler_step(tsp_solver *s) {
    if (!s->launch) return fail(TSP_UNSUPPORTED, "a custom model steps through tsp_euler_steps_custom");
    int &OLD = s->role[TSP_ROLE_OLD], &NEW = s->role[TSP_ROLE_NEW];
    if (fills_ghosts(s))
        if (tsp_status f = fill_patch(s, OLD, s->time); f != TSP_OK) return f;
    const tsp::StageArgs a = stage_args(s, tsp::kEuler, OLD, OLD, NEW, -1, s->d.dt, 0.0);
    TSP_HIP(s->launch(a, s->cfg.tile_y, s->stream));
    std::swap(OLD, NEW);
    return TSP_OK;
}

void advance_clock(tsp_solver *s) { // Integrator.hpp:98-100
    s->steps++;
    s->time += s->d.dt;
}

// The captured two-step graph for register roles `roles` (capturing it if
// missing); graphs of another kernel configuration are dropped first.
tsp_status ensure_graph(tsp_solver *s, const int roles[4], int *slot) {
    if (s->d.bc == TSP_BC_CUSTOM)
        return fail(TSP_UNSUPPORTED, "a custom boundary functor is called with each fill's time; graph replay "
                                     "would freeze those times (step eagerly)");
    if (std::memcmp(&s->graph_cfg, &s->cfg, sizeof(tsp_kernel_cfg)) != 0) {
        drop_graphs(s);
        s->graph_cfg = s->cfg;
    }
    for (int g = 0; g < 2; ++g)
        if (s->graph[g] && std::memcmp(s->graph_role[g], roles, sizeof(s->graph_role[g])) == 0) {
            *slot = g;
            return TSP_OK;
        }
    const int g = !s->graph[0] ? 0 : !s->graph[1] ? 1 : s->graph_next;
    s->graph_next = 1 - g;
    if (s->graph[g]) {
        (void)hipGraphExecDestroy(s->graph[g]);
        s->graph[g] = nullptr;
    }
    int saved[4];
    std::memcpy(saved, s->role, sizeof(saved));
    std::memcpy(s->role, roles, sizeof(saved));
    hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        std::memcpy(s->role, saved, sizeof(saved));
        return hip_fail(e, "hipStreamBeginCapture");
    }
    tsp_status st = enqueue_rk4_step(s, nullptr);
    if (st == TSP_OK) st = enqueue_rk4_step(s, nullptr);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s->stream, &graph);
    std::memcpy(s->role, saved, sizeof(saved));
    if (st != TSP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&s->graph[g], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        s->graph[g] = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    std::memcpy(s->graph_role[g], roles, sizeof(s->graph_role[g]));
    *slot = g;
    return TSP_OK;
}

} // namespace

extern "C" {

const char *tsp_last_error(void) { return g_err.c_str(); }
const char *tsp_version(void) { return TSP_VERSION; }

int64_t tsp_rk4_bytes_per_cell(const tsp_desc *d) {
    const int nc = d ? desc_ncomp(*d) : -1;
    return nc < 0 ? -1 : 128LL * nc;
}

tsp_status tsp_desc_default(tsp_desc *d, int32_t system, int32_t n) {
    if (!d) return fail(TSP_INVALID_ARG, "descriptor is NULL");
    if (system != TSP_SYS_CUSTOM && system_ncomp(system) < 0) return fail(TSP_INVALID_ARG, "unknown system id");
    std::memset(d, 0, sizeof(*d));
    d->system = system;
    if (system == TSP_SYS_CUSTOM) d->ncomp = 1;
    for (int a = 0; a < 3; ++a) {
        d->n[a] = n;
        d->prob_lo[a] = 0.0;
        d->prob_hi[a] = 1.0;
        d->periodic[a] = 1;
    }
    d->stencil_n = 2;
    d->ghost = 2;
    d->ko_order = 0;
    d->bc = TSP_BC_NOOP;
    d->abl_profile = TSP_PROFILE_QUADRATIC;
    d->dt = 0.25 * (1.0 / n);
    return TSP_OK;
}

tsp_status tsp_create(const tsp_desc *d, void *stream, tsp_solver **out) {
    if (!out) return fail(TSP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    tsp_status st = validate(d);
    if (st != TSP_OK) return st;
    if (d->system != TSP_SYS_TENSOR_ADM && (st = check_abl_lds(d->abl_thickness)) != TSP_OK) return st;
    tsp_solver *s = new (std::nothrow) tsp_solver();
    if (!s) return fail(TSP_OOM, "host allocation failed");
    s->d = *d;
    s->nc = desc_ncomp(*d);
    s->G = d->ghost;
    s->H = std::max(d->stencil_n, d->ko_order);
    s->pitch = (int)(((int64_t)kXOff + d->n[0] + s->G + 15) / 16 * 16);
    s->sj = s->pitch;
    s->sk = s->sj * (d->n[1] + 2 * s->G);
    s->sc = s->sk * (d->n[2] + 2 * s->G);
    s->reg_elems = s->sc * s->nc;
    s->launch = lookup_launcher(d->system, d->stencil_n, d->ko_order);
    for (int st = 0; st < 4; ++st) s->launch_stage[st] = lookup_stage_launcher(d->system, d->stencil_n, d->ko_order, st);
    if (d->system == TSP_SYS_WAVE_PHIPI) s->launch_pair = tsp::lookup_pair_phipi(d->stencil_n, d->ko_order);
    s->cfg = default_cfg(*d);
    if (hipGetDevice(&s->device) != hipSuccess) s->device = -1;
    s->ptr_access = std::getenv("TSP_POINTER_ACCESS") != nullptr;
    // Geometry::CellSize = (prob_hi - prob_lo) / n_cells; grid factors as the stencils form them
    for (int a = 0; a < 3; ++a) {
        const double dx = (d->prob_hi[a] - d->prob_lo[a]) / (double)d->n[a];
        s->gf1[a] = 1.0 / dx;
        s->gf2[a] = 1.0 / (dx * dx);
    }
    auto cleanup = [&](tsp_status e) {
        tsp_destroy(s);
        return e;
    };
    if (stream) {
        s->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipStreamCreateWithFlags"));
        s->own_stream = true;
    }
    for (int r = 0; r < 4; ++r) {
        hipError_t e = hipMalloc((void **)&s->base[r], sizeof(double) * s->reg_elems);
        if (e == hipErrorOutOfMemory) return cleanup(fail(TSP_OOM, "hipMalloc of a register failed (out of memory)"));
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc"));
        e = hipMemsetAsync(s->base[r], 0, sizeof(double) * s->reg_elems, s->stream);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMemsetAsync"));
    }
    if (d->system == TSP_SYS_TENSOR_ADM && d->ko_order > 0) {
        const size_t bytes = sizeof(double) * 6 * (size_t)s->sc;
        hipError_t e = hipMalloc((void **)&s->kko, bytes);
        if (e == hipErrorOutOfMemory) return cleanup(fail(TSP_OOM, "hipMalloc of the KO buffer failed (out of memory)"));
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc"));
        e = hipMemsetAsync(s->kko, 0, bytes, s->stream);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMemsetAsync"));
    }
    // Ghost cells outside a non-periodic face only ever hold the boundary
    // values (stages write valid cells only), so they are written once here
    // for every register; the virtual-ghost stage kernels read them as is.
    for (int r = 0; r < 4; ++r) {
        hipError_t e = tsp::launch_ghost_fill(ghost_args(s, r), s->stream);
        if (e != hipSuccess) return cleanup(hip_fail(e, "ghost prefill"));
    }
    if (d->abl_thickness > 0) {
        // ABLDecorator.hpp:115-132 tabulated by squared integer distance to the bulk box
        const int T = d->abl_thickness;
        const int nt = 3 * T * T + 1;
        std::vector<double> tab(nt, 0.0);
        const double inv_thickness = 1.0 / static_cast<double>(T);
        for (int d2 = 1; d2 < nt; ++d2) {
            const double norm = std::sqrt(static_cast<double>(d2));
            const double x = inv_thickness * norm;
            const double scaled = (x < 0.0) ? 0.0 : (1.0 < x) ? 1.0 : x;
            tab[d2] = d->abl_max * abl_profile(d->abl_profile, scaled);
        }
        hipError_t e = hipMalloc((void **)&s->sigma_tab, sizeof(double) * nt);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(sigma table)"));
        e = hipMemcpy(s->sigma_tab, tab.data(), sizeof(double) * nt, hipMemcpyHostToDevice);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMemcpy(sigma table)"));
    }
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipStreamSynchronize"));
    *out = s;
    return TSP_OK;
}

tsp_status tsp_destroy(tsp_solver *s) {
    if (!s) return TSP_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    for (int r = 0; r < 4; ++r)
        if (s->base[r]) (void)hipFree(s->base[r]);
    if (s->sigma_tab) (void)hipFree(s->sigma_tab);
    if (s->kko) (void)hipFree(s->kko);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return TSP_OK;
}

tsp_status tsp_set_state(tsp_solver *s, const double *src, int32_t on_device) {
    if (!s || !src) return fail(TSP_INVALID_ARG, "solver or source is NULL");
    const int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    const size_t bytes = sizeof(double) * (size_t)nx * ny * nz * s->nc;
    const double *dsrc = src;
    double *staging = nullptr;
    if (!on_device) {
        TSP_HIP(hipMalloc((void **)&staging, bytes));
        hipError_t e = hipMemcpyAsync(staging, src, bytes, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) {
            (void)hipFree(staging);
            return hip_fail(e, "hipMemcpyAsync(state)");
        }
        dsrc = staging;
    }
    // The register roles are a function of the step parity (captured graphs
    // rely on it), so restarting the clock restarts the roles too.
    for (int r = 0; r < 4; ++r) s->role[r] = r;
    const int OLD = s->role[TSP_ROLE_OLD];
    hipError_t e = tsp::launch_unpack(reg_origin(s, OLD), dsrc, nx, ny, nz, s->nc, (int)s->sj, (int)s->sk, (int)s->sc,
                                      s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (staging) (void)hipFree(staging);
    if (e != hipSuccess) return hip_fail(e, "set_state");
    s->time = 0.0;
    s->steps = 0;
    return TSP_OK;
}

tsp_status tsp_get_state(tsp_solver *s, double *dst, int32_t on_device) {
    if (!s || !dst) return fail(TSP_INVALID_ARG, "solver or destination is NULL");
    const int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    const size_t bytes = sizeof(double) * (size_t)nx * ny * nz * s->nc;
    double *ddst = dst;
    double *staging = nullptr;
    if (!on_device) {
        TSP_HIP(hipMalloc((void **)&staging, bytes));
        ddst = staging;
    }
    const int OLD = s->role[TSP_ROLE_OLD];
    hipError_t e =
        tsp::launch_pack(ddst, reg_origin(s, OLD), nx, ny, nz, s->nc, (int)s->sj, (int)s->sk, (int)s->sc, s->stream);
    if (e == hipSuccess && staging) e = hipMemcpyAsync(dst, staging, bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (staging) (void)hipFree(staging);
    if (e != hipSuccess) return hip_fail(e, "get_state");
    return TSP_OK;
}

tsp_status tsp_init_expr(tsp_solver *s, const tsp_expr_op *ops, const int32_t *comp_start, const int32_t index_lo[3]) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (tsp_status st = check_expr_programs(ops, comp_start, s->nc); st != TSP_OK) return st;
    for (int r = 0; r < 4; ++r) s->role[r] = r; // as tsp_set_state: the clock and the roles restart
    tsp::InitBox box{reg_origin(s, s->role[TSP_ROLE_OLD]), s->sj, s->sk, s->sc, {s->d.n[0], s->d.n[1], s->d.n[2]},
                     {0, 0, 0}};
    if (tsp_status st = run_init_expr({box}, s->d, ops, comp_start, s->nc, index_lo, s->stream); st != TSP_OK)
        return st;
    s->time = 0.0;
    s->steps = 0;
    return TSP_OK;
}

tsp_status tsp_fill_ghosts(tsp_solver *s, int32_t role) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (role < 0 || role > 3) return fail(TSP_INVALID_ARG, "role must be 0..3");
    return fill_patch(s, s->role[role], s->time);
}

tsp_status tsp_set_bc_callback(tsp_solver *s, tsp_bc_fn fn, void *user) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (s->d.bc != TSP_BC_CUSTOM) return fail(TSP_INVALID_ARG, "the solver's boundary functor is not TSP_BC_CUSTOM");
    s->bc_fn = fn;
    s->bc_user = user;
    return TSP_OK;
}

tsp_status tsp_get_register(tsp_solver *s, int32_t role, double *dst) {
    if (!s || !dst) return fail(TSP_INVALID_ARG, "solver or destination is NULL");
    if (role < 0 || role > 3) return fail(TSP_INVALID_ARG, "role must be 0..3");
    const int G = s->G;
    const int gx = s->d.n[0] + 2 * G;
    const size_t rows = (size_t)(s->d.n[1] + 2 * G) * (s->d.n[2] + 2 * G) * s->nc;
    const double *src = s->base[s->role[role]] + kXOff - G;
    TSP_HIP(hipStreamSynchronize(s->stream));
    TSP_HIP(hipMemcpy2D(dst, sizeof(double) * gx, src, sizeof(double) * s->pitch, sizeof(double) * gx, rows,
                        hipMemcpyDeviceToHost));
    return TSP_OK;
}

// ---------------------------------------------------------------------------
// RK4 steps without any allocation: the four registers in a file-scope device
// array (one copy per device, shared by every call on that device)
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = size_t(128) << 20; // TSP_NOALLOC_SCRATCH_BYTES
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up once per device
tsp_status scratch_of_current_device(double **out) {
    int dev = -1;
    TSP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TSP_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return TSP_OK;
}
} // namespace

tsp_status tsp_rk4_steps_noalloc(tsp_solver *s, const double *state_in, double *state_out, int64_t nsteps) {
    if (!s || !state_in || !state_out) return fail(TSP_INVALID_ARG, "solver, input or output is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    if (!s->launch) return fail(TSP_UNSUPPORTED, "a custom model steps through tsp_rk4_steps_custom");
    const int64_t reg = (s->reg_elems + 31) / 32 * 32; // 256-byte aligned registers
    if (4 * reg * (int64_t)sizeof(double) > (int64_t)kScratchBytes)
        return fail(TSP_INVALID_ARG, "the four registers of this solver (" +
                                         std::to_string(4 * reg * (int64_t)sizeof(double)) +
                                         " bytes) exceed the " + std::to_string(kScratchBytes) +
                                         "-byte scratch of tsp_rk4_steps_noalloc");
    int dev = -1;
    TSP_HIP(hipGetDevice(&dev));
    if (dev != s->device) return fail(TSP_INVALID_ARG, "the solver was created on another device than the current one");
    double *ptr = nullptr;
    tsp_status st = scratch_of_current_device(&ptr);
    if (st != TSP_OK) return st;
    // a stack copy of the solver stepping the scratch registers: same descriptor,
    // kernels and configuration; the solver's own registers, roles and clock are untouched
    tsp_solver t = *s;
    t.graph[0] = t.graph[1] = nullptr;
    for (int r = 0; r < 4; ++r) {
        t.base[r] = ptr + r * reg;
        t.role[r] = r;
    }
    // the scratch holds the last call's contents: zero every register (the NoOp
    // ghosts are read as zero, as in a new solver), write the Dirichlet ghosts,
    // then the caller's state
    for (int r = 0; r < 4; ++r) {
        TSP_HIP(tsp::launch_zero(t.base[r], s->reg_elems, s->stream));
        TSP_HIP(tsp::launch_ghost_fill(ghost_args(&t, r), s->stream));
    }
    TSP_HIP(tsp::launch_unpack(reg_origin(&t, t.role[TSP_ROLE_OLD]), state_in, s->d.n[0], s->d.n[1], s->d.n[2], s->nc,
                               (int)s->sj, (int)s->sk, (int)s->sc, s->stream));
    for (int64_t n = 0; n < nsteps; ++n) {
        st = enqueue_rk4_step(&t, nullptr);
        if (st != TSP_OK) return st;
        advance_clock(&t); // the copy's clock: the fill times of a custom boundary functor
    }
    TSP_HIP(tsp::launch_pack(state_out, reg_origin(&t, t.role[TSP_ROLE_OLD]), s->d.n[0], s->d.n[1], s->d.n[2], s->nc,
                             (int)s->sj, (int)s->sk, (int)s->sc, s->stream));
    return TSP_OK;
}

tsp_status tsp_rk4_steps(tsp_solver *s, int64_t nsteps) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    for (int64_t n = 0; n < nsteps; ++n) {
        tsp_status st = enqueue_rk4_step(s, nullptr);
        if (st != TSP_OK) return st;
        advance_clock(s);
    }
    return TSP_OK;
}

tsp_status tsp_euler_steps(tsp_solver *s, int64_t nsteps) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    for (int64_t n = 0; n < nsteps; ++n) {
        tsp_status st = enqueue_euler_step(s);
        if (st != TSP_OK) return st;
        advance_clock(s);
    }
    return TSP_OK;
}

namespace {
tsp_stage_view stage_view(const tsp_solver *s, int stage, int src, int init, int acc, int argn, double time, double b,
                          double a_next) {
    tsp_stage_view v{};
    v.stage = stage;
    for (int a = 0; a < 3; ++a) v.n[a] = s->d.n[a];
    v.ghost = s->G;
    v.ncomp = s->nc;
    v.time = time;
    v.b = b;
    v.a_next = a_next;
    v.src = reg_origin(s, src);
    v.init = reg_origin(s, init);
    v.acc = reg_origin(s, acc);
    v.argn = argn >= 0 ? reg_origin(s, argn) : nullptr;
    v.sj = s->sj;
    v.sk = s->sk;
    v.sc = s->sc;
    return v;
}

tsp_status call_stage(tsp_solver *s, tsp_stage_fn fn, void *user, const tsp_stage_view &v) {
    const int r = fn(user, &v, (void *)s->stream);
    if (r != 0) return fail(TSP_HIP_ERROR, "custom stage kernel failed (callback returned " + std::to_string(r) + ")");
    return TSP_OK;
}
} // namespace

tsp_status tsp_rk4_steps_custom(tsp_solver *s, int64_t nsteps, tsp_stage_fn fn, void *user) {
    if (!s || !fn) return fail(TSP_INVALID_ARG, "solver or stage callback is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    const double dt = s->d.dt;
    int &OLD = s->role[TSP_ROLE_OLD], &NEW = s->role[TSP_ROLE_NEW], &ARG = s->role[TSP_ROLE_ARG],
        &ARGN = s->role[TSP_ROLE_ARG_NEXT];
    for (int64_t n = 0; n < nsteps; ++n) {
        if (fills_ghosts(s))
            if (tsp_status f = fill_patch(s, OLD, s->time); f != TSP_OK) return f; // pre_step FillPatch
        for (int stage = 0; stage < 4; ++stage) {
            // integrators/SingleLevelRK4Integrator.hpp:115-125
            const double a_next = dt * kAFrac[(stage + 1) % 4];
            const double time_arg = s->time + dt * kAFrac[stage];
            const double b = dt * kBFrac[stage];
            if (stage > 0) {
                std::swap(ARG, ARGN);
                if (fills_ghosts(s))
                    if (tsp_status f = fill_patch(s, ARG, time_arg); f != TSP_OK) return f;
            }
            const tsp_stage_view v =
                stage_view(s, stage, stage == 0 ? OLD : ARG, OLD, NEW, stage < 3 ? ARGN : -1, time_arg, b, a_next);
            tsp_status st = call_stage(s, fn, user, v);
            if (st != TSP_OK) return st;
        }
        std::swap(OLD, NEW); // post_step
        advance_clock(s);
    }
    return TSP_OK;
}

tsp_status tsp_euler_steps_custom(tsp_solver *s, int64_t nsteps, tsp_stage_fn fn, void *user) {
    if (!s || !fn) return fail(TSP_INVALID_ARG, "solver or stage callback is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    int &OLD = s->role[TSP_ROLE_OLD], &NEW = s->role[TSP_ROLE_NEW];
    for (int64_t n = 0; n < nsteps; ++n) {
        if (fills_ghosts(s))
            if (tsp_status f = fill_patch(s, OLD, s->time); f != TSP_OK) return f;
        const tsp_stage_view v = stage_view(s, tsp::kEuler, OLD, OLD, NEW, -1, s->time, s->d.dt, 0.0);
        tsp_status st = call_stage(s, fn, user, v);
        if (st != TSP_OK) return st;
        std::swap(OLD, NEW);
        advance_clock(s);
    }
    return TSP_OK;
}

tsp_status tsp_get_time(tsp_solver *s, double *time, int64_t *steps) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (time) *time = s->time;
    if (steps) *steps = s->steps;
    return TSP_OK;
}

tsp_status tsp_get_kernel(tsp_solver *s, tsp_kernel_cfg *cfg) {
    if (!s || !cfg) return fail(TSP_INVALID_ARG, "solver or cfg is NULL");
    *cfg = s->cfg;
    return TSP_OK;
}

tsp_status tsp_set_kernel(tsp_solver *s, const tsp_kernel_cfg *cfg) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    tsp_status st = validate_cfg(cfg);
    if (st != TSP_OK) return st;
    std::string why;
    if (cfg->stage_pairs && !pairs_apply(s, cfg->ghost_mode, &why)) return fail(TSP_UNSUPPORTED, why);
    if (s->d.bc == TSP_BC_CUSTOM && cfg->ghost_mode != 1)
        return fail(TSP_INVALID_ARG, "a custom boundary functor needs ghost_mode 1 (the fill pass)");
    s->cfg = *cfg;
    return TSP_OK;
}

tsp_status tsp_sync(tsp_solver *s) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    TSP_HIP(hipStreamSynchronize(s->stream));
    return TSP_OK;
}

tsp_status tsp_get_stream(tsp_solver *s, void **stream) {
    if (!s || !stream) return fail(TSP_INVALID_ARG, "NULL argument");
    *stream = (void *)s->stream;
    return TSP_OK;
}

// ---------------------------------------------------------------------------
// device scope timer (ScopedTimer.hpp:34-149 measures the host clock)
// ---------------------------------------------------------------------------

struct tsp_timer {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    bool running = false;
};

tsp_status tsp_timer_create(void *stream, tsp_timer **out) {
    if (!out) return fail(TSP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    tsp_timer *t = new (std::nothrow) tsp_timer;
    if (!t) return fail(TSP_OOM, "host allocation failed");
    t->stream = (hipStream_t)stream;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) {
        t->start = ev[0];
        e = hipEventCreate(&ev[1]);
        if (e == hipSuccess) t->stop = ev[1];
    }
    if (e != hipSuccess) {
        tsp_timer_destroy(t);
        return fail(TSP_HIP_ERROR, std::string("hipEventCreate: ") + hipGetErrorString(e));
    }
    *out = t;
    return TSP_OK;
}

tsp_status tsp_timer_start(tsp_timer *t) {
    if (!t) return fail(TSP_INVALID_ARG, "timer is NULL");
    TSP_HIP(hipEventRecord(t->start, t->stream));
    t->running = true;
    return TSP_OK;
}

tsp_status tsp_timer_stop(tsp_timer *t, double *elapsed_ms) {
    if (!t || !elapsed_ms) return fail(TSP_INVALID_ARG, "NULL argument");
    if (!t->running) return fail(TSP_INVALID_ARG, "timer was not started");
    t->running = false;
    TSP_HIP(hipEventRecord(t->stop, t->stream));
    TSP_HIP(hipEventSynchronize(t->stop));
    float ms = 0.0f;
    TSP_HIP(hipEventElapsedTime(&ms, t->start, t->stop));
    *elapsed_ms = ms;
    return TSP_OK;
}

tsp_status tsp_timer_destroy(tsp_timer *t) {
    if (!t) return TSP_OK;
    if (t->start) (void)hipEventDestroy(t->start);
    if (t->stop) (void)hipEventDestroy(t->stop);
    delete t;
    return TSP_OK;
}

tsp_status tsp_register_view(tsp_solver *s, int32_t role, double **ptr, int64_t *sj, int64_t *sk, int64_t *sc) {
    if (!s || !ptr) return fail(TSP_INVALID_ARG, "solver or ptr is NULL");
    if (role < 0 || role > 3) return fail(TSP_INVALID_ARG, "role must be 0..3");
    *ptr = reg_origin(s, s->role[role]);
    if (sj) *sj = s->sj;
    if (sk) *sk = s->sk;
    if (sc) *sc = s->sc;
    return TSP_OK;
}

tsp_status tsp_graph_rk4_steps(tsp_solver *s, int64_t nsteps) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    if (s->d.bc == TSP_BC_CUSTOM)
        return fail(TSP_UNSUPPORTED, "a custom boundary functor is called with each fill's time; graph replay "
                                     "would freeze those times (step eagerly)");
    int slot = -1;
    if (nsteps >= 2) {
        tsp_status st = ensure_graph(s, s->role, &slot);
        if (st != TSP_OK) return st;
    }
    int64_t n = 0;
    for (; n + 2 <= nsteps; n += 2) {
        TSP_HIP(hipGraphLaunch(s->graph[slot], s->stream));
        advance_clock(s);
        advance_clock(s);
    }
    if (n < nsteps) return tsp_rk4_steps(s, nsteps - n);
    return TSP_OK;
}

tsp_status tsp_graph_prepare(tsp_solver *s) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    int slot = -1;
    tsp_status st = ensure_graph(s, s->role, &slot);
    if (st != TSP_OK) return st;
    // roles one RK4 step later: old <-> new, arg <-> arg_next
    const int next[4] = {s->role[TSP_ROLE_NEW], s->role[TSP_ROLE_OLD], s->role[TSP_ROLE_ARG_NEXT],
                         s->role[TSP_ROLE_ARG]};
    return ensure_graph(s, next, &slot);
}

tsp_status tsp_time_rk4(tsp_solver *s, int64_t nsteps, int32_t use_graph, double *ms) {
    if (!s || !ms) return fail(TSP_INVALID_ARG, "solver or ms is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TSP_HIP(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "hipEventCreate");
    }
    tsp_status st = TSP_OK;
    if ((e = hipEventRecord(e0, s->stream)) != hipSuccess) st = hip_fail(e, "hipEventRecord");
    if (st == TSP_OK) st = use_graph ? tsp_graph_rk4_steps(s, nsteps) : tsp_rk4_steps(s, nsteps);
    if (st == TSP_OK && (e = hipEventRecord(e1, s->stream)) != hipSuccess) st = hip_fail(e, "hipEventRecord");
    if (st == TSP_OK && (e = hipEventSynchronize(e1)) != hipSuccess) st = hip_fail(e, "hipEventSynchronize");
    float f = 0.f;
    if (st == TSP_OK && (e = hipEventElapsedTime(&f, e0, e1)) != hipSuccess) st = hip_fail(e, "hipEventElapsedTime");
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

tsp_status tsp_profile_rk4(tsp_solver *s, int64_t nsteps, double *stage_ms, double *fill_ms) {
    if (!s || !stage_ms) return fail(TSP_INVALID_ARG, "solver or stage_ms is NULL");
    if (nsteps < 1) return fail(TSP_INVALID_ARG, "nsteps must be >= 1");
    for (int i = 0; i < 4; ++i) stage_ms[i] = 0.0;
    if (fill_ms) *fill_ms = 0.0;
    StepTiming tm;
    tsp_status st = TSP_OK;
    for (EventPair *p : {&tm.stage[0], &tm.stage[1], &tm.stage[2], &tm.stage[3], &tm.fill[0], &tm.fill[1],
                         &tm.fill[2], &tm.fill[3]}) {
        hipError_t e = hipEventCreate(&p->start);
        if (e == hipSuccess) e = hipEventCreate(&p->stop);
        if (e != hipSuccess) st = hip_fail(e, "hipEventCreate");
    }
    for (int64_t n = 0; n < nsteps && st == TSP_OK; ++n) {
        st = enqueue_rk4_step(s, &tm);
        if (st != TSP_OK) break;
        advance_clock(s);
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            st = hip_fail(e, "hipStreamSynchronize");
            break;
        }
        for (int i = 0; i < 4 && st == TSP_OK; ++i) {
            float ms = 0.f;
            e = hipEventElapsedTime(&ms, tm.stage[i].start, tm.stage[i].stop);
            if (e != hipSuccess) st = hip_fail(e, "hipEventElapsedTime");
            stage_ms[i] += ms;
            if (fills_ghosts(s) && fill_ms && st == TSP_OK) {
                e = hipEventElapsedTime(&ms, tm.fill[i].start, tm.fill[i].stop);
                if (e != hipSuccess) st = hip_fail(e, "hipEventElapsedTime");
                *fill_ms += ms;
            }
        }
    }
    for (EventPair *p : {&tm.stage[0], &tm.stage[1], &tm.stage[2], &tm.stage[3], &tm.fill[0], &tm.fill[1],
                         &tm.fill[2], &tm.fill[3]}) {
        if (p->start) (void)hipEventDestroy(p->start);
        if (p->stop) (void)hipEventDestroy(p->stop);
    }
    return st;
}

namespace {

std::string tune_key(const tsp_solver *s) {
    hipDeviceProp_t prop{};
    int dev = 0;
    std::string arch = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
    for (char &c : arch)
        if (c == ' ' || c == ':') c = '_';
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s|sys%d|n%d|r%d|%dx%dx%d|abl%d", arch.c_str(), s->d.system, s->d.stencil_n,
                  s->d.ko_order, s->d.n[0], s->d.n[1], s->d.n[2], s->d.abl_thickness > 0 ? 1 : 0);
    return buf;
}

// Number of whitespace-separated fields of a cache line.
int count_fields(const char *line) {
    int n = 0;
    bool in = false;
    for (const char *p = line; *p; ++p) {
        const bool sp = *p == ' ' || *p == '\t' || *p == '\n' || *p == '\r';
        if (!sp && !in) ++n;
        in = !sp;
    }
    return n;
}

// A cached configuration for `key` that this solver can run (the last matching line wins).
// Lines are "key ghost_mode tile_y zchunk xcd_remap zalt halo_ring stage_pairs ms"; lines with
// another field count (the format before stage_pairs) are skipped.
bool cache_lookup(const tsp_solver *s, const std::string &key, tsp_kernel_cfg *cfg) {
    const char *path = std::getenv("TSP_TUNE_CACHE");
    if (!path || !*path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        char k[256];
        tsp_kernel_cfg c{};
        double ms = 0;
        if (count_fields(line) == 9 &&
            std::sscanf(line, "%255s %d %d %d %d %d %d %d %lf", k, &c.ghost_mode, &c.tile_y, &c.zchunk, &c.xcd_remap,
                        &c.zalt, &c.halo_ring, &c.stage_pairs, &ms) == 9 &&
            key == k && validate_cfg(&c) == TSP_OK && (!c.stage_pairs || pairs_apply(s, c.ghost_mode, nullptr)) &&
            (s->d.bc != TSP_BC_CUSTOM || c.ghost_mode == 1)) {
            *cfg = c;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

void cache_store(const std::string &key, const tsp_kernel_cfg &c, double ms) {
    const char *path = std::getenv("TSP_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d %d %d %d %d %.6f\n", key.c_str(), c.ghost_mode, c.tile_y, c.zchunk, c.xcd_remap,
                 c.zalt, c.halo_ring, c.stage_pairs, ms);
    std::fclose(f);
}

} // namespace

tsp_status tsp_autotune(tsp_solver *s, int32_t steps, tsp_kernel_cfg *chosen) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (steps < 1) steps = 1;
    const std::string key = tune_key(s);
    tsp_kernel_cfg best = s->cfg;
    if (s->d.system == TSP_SYS_CUSTOM) return fail(TSP_UNSUPPORTED, "a custom model has no library stage kernel to tune");
    if (cache_lookup(s, key, &best)) {
        s->cfg = best;
        if (chosen) *chosen = best;
        return TSP_OK;
    }
    // save the state: roles, clock and the contents of state_old
    int saved_role[4];
    std::memcpy(saved_role, s->role, sizeof(saved_role));
    const double saved_time = s->time;
    const int64_t saved_steps = s->steps;
    const tsp_kernel_cfg saved_cfg = s->cfg;
    double *backup = nullptr;
    TSP_HIP(hipMalloc((void **)&backup, sizeof(double) * s->reg_elems));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    tsp_status st = TSP_OK;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = hipMemcpyAsync(backup, s->base[saved_role[TSP_ROLE_OLD]], sizeof(double) * s->reg_elems,
                           hipMemcpyDeviceToDevice, s->stream);
    if (e != hipSuccess) st = hip_fail(e, "autotune setup");

    std::vector<int> zchunks;
    for (int z : {8, 16, 32, 64, 128, 256})
        if (z < s->d.n[2]) zchunks.push_back(z);
    zchunks.push_back(s->d.n[2]);
    // candidates: every tile height, z chunk, XCD mapping, z-march order and
    // (for the tile heights that have one) halo ring with virtual ghosts; the
    // fastest four are timed again over more steps, then the best geometry
    // with the fill-pass mode
    std::vector<tsp_kernel_cfg> cands;
    const bool tensor = s->d.system == TSP_SYS_TENSOR_ADM;
    if (tensor) { // block shape (tile_y) and z tiles per slab of the block order (zchunk)
        for (int ty : {8, 4, 16})
            for (int zc : {1, 2, 4}) cands.push_back(tsp_kernel_cfg{1, ty, zc, 1, 0, 0, 0});
    } else {
        // a custom boundary functor needs the fill pass (ghost_mode 1) in every candidate
        const int gm = s->d.bc == TSP_BC_CUSTOM ? 1 : 0;
        for (int ty : {2, 4, 8, 16, 32})
            for (int zc : zchunks)
                for (int xr = 0; xr < 2; ++xr)
                    for (int za = 0; za < 3; ++za)
                        for (int hr = 0; hr < (ty >= 8 ? 2 : 1); ++hr)
                            cands.push_back(tsp_kernel_cfg{gm, ty, zc, xr, za, hr, 0});
    }
    // stage pairs (32 x 32 tiles; tile_y, zalt and halo_ring do not apply)
    if (pairs_apply(s, 0, nullptr))
        for (int zc : zchunks)
            for (int xr = 0; xr < 2; ++xr) cands.push_back(tsp_kernel_cfg{0, s->cfg.tile_y, zc, xr, 0, 0, 1});
    double best_ms = 1e300;
    std::vector<std::pair<double, tsp_kernel_cfg>> timed;
    auto time_cfg = [&](const tsp_kernel_cfg &c, int nsteps, double *out) -> tsp_status {
        s->cfg = c;
        tsp_status r = enqueue_rk4_step(s, nullptr); // warm-up
        if (r != TSP_OK) return r;
        if ((e = hipEventRecord(e0, s->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        for (int n = 0; n < nsteps; ++n)
            if ((r = enqueue_rk4_step(s, nullptr)) != TSP_OK) return r;
        if ((e = hipEventRecord(e1, s->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((e = hipEventSynchronize(e1)) != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
        *out = ms / nsteps;
        // restart every candidate from the saved state so values stay bounded
        std::memcpy(s->role, saved_role, sizeof(saved_role));
        e = hipMemcpyAsync(s->base[saved_role[TSP_ROLE_OLD]], backup, sizeof(double) * s->reg_elems,
                           hipMemcpyDeviceToDevice, s->stream);
        return e == hipSuccess ? TSP_OK : hip_fail(e, "autotune restore");
    };
    for (size_t i = 0; i < cands.size() && st == TSP_OK; ++i) {
        double per = 0.0;
        st = time_cfg(cands[i], steps, &per);
        if (st == TSP_OK) timed.emplace_back(per, cands[i]);
    }
    if (st == TSP_OK && !timed.empty()) {
        // the fastest few again over five times the steps, twice, keeping each one's better
        // time; within 0.5 % of the fastest, the halo ring (fewer HBM bytes) is preferred
        std::stable_sort(timed.begin(), timed.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        const size_t top = std::min<size_t>(timed.size(), 4);
        std::vector<double> fine(top, 1e300);
        for (int rep = 0; rep < 2 && st == TSP_OK; ++rep)
            for (size_t i = 0; i < top && st == TSP_OK; ++i) {
                double per = 0.0;
                st = time_cfg(timed[i].second, 5 * steps, &per);
                fine[i] = std::min(fine[i], per);
            }
        if (st == TSP_OK) {
            size_t bi = 0;
            for (size_t i = 1; i < top; ++i)
                if (fine[i] < fine[bi]) bi = i;
            for (size_t i = 0; i < top; ++i)
                if (fine[i] <= fine[bi] * 1.005 && timed[i].second.halo_ring > timed[bi].second.halo_ring) bi = i;
            best = timed[bi].second;
            best_ms = fine[bi];
        }
    }
    if (st == TSP_OK && !tensor) { // (the tensor kernel has no virtual-ghost mode to compare with)
        tsp_kernel_cfg fill = best;
        fill.ghost_mode = 1;
        fill.stage_pairs = 0;
        double per = 0.0;
        st = time_cfg(fill, 5 * steps, &per);
        if (st == TSP_OK && per < best_ms) {
            best_ms = per;
            best = fill;
        }
    }
    // restore
    std::memcpy(s->role, saved_role, sizeof(saved_role));
    e = hipMemcpyAsync(s->base[saved_role[TSP_ROLE_OLD]], backup, sizeof(double) * s->reg_elems,
                       hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess && st == TSP_OK) st = hip_fail(e, "autotune restore");
    (void)hipFree(backup);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    s->time = saved_time;
    s->steps = saved_steps;
    if (st != TSP_OK) {
        s->cfg = saved_cfg;
        return st;
    }
    s->cfg = best;
    cache_store(key, best, best_ms);
    if (chosen) *chosen = best;
    return TSP_OK;
}

} // extern "C"

// ===========================================================================
// Discrete adjoint of the RK4 stepping for the linear (phi, Pi) system
// ===========================================================================
//
// J = sum_{n=1..N} j_n(phi^n at the receivers). With G the RK4 step map
// (affine; its linear part is the same G for every step) the gradient with
// respect to the initial state is obtained by
//   lambda = 0;  for n = N..1:  lambda += dj_n/du^n;  lambda = G^T lambda
// where G^T is applied by the reverse-stage recursion
//   mu_3 = b_3 lambda';  nu_j = L^T mu_j;  mu_{j-1} = b_{j-1} lambda' + a_j nu_j;
//   lambda = lambda' + nu_3 + nu_2 + nu_1 + nu_0
// (b_j = dt * b_fraction[j], a_j = dt * a_fraction[j],
// integrators/SingleLevelRK4Integrator.hpp:71-74), each reverse stage one
// launch of the fused stage kernel on the transpose operator.

struct tsp_adjoint {
    tsp_solver *s = nullptr;
    int nrec = 0, nsteps = 0, K = 0, nck = 0;
    int64_t step0 = 0;  // solver step count at the start of the forward pass
    double time0 = 0.0; // and its clock
    int *d_rec = nullptr;     // receiver offsets of component 0 relative to the register origin
    double *d_trace = nullptr;
    double *d_dtrace = nullptr;
    int *d_counter = nullptr;
    double *d_ckpt = nullptr; // nck dense states: steps 0, K, 2K, ... and the final step
    int64_t state_elems = 0;
    double *abase[4] = {nullptr, nullptr, nullptr, nullptr};
    int arole[4] = {0, 1, 2, 3};
    int roles0[4] = {0, 1, 2, 3};      // the solver's register roles when the forward pass started
    int groles0[4] = {-1, -1, -1, -1}; // roles0 the forward graphs were captured for
    tsp::LaunchFn launch = nullptr;
    hipGraphExec_t fgraph[2] = {nullptr, nullptr}; // two (step, gather) pairs per forward parity
    hipGraphExec_t bgraph[2] = {nullptr, nullptr}; // two (inject, reverse step) pairs per backward parity
    tsp_kernel_cfg gcfg{-1, -1, -1, -1, -1, -1, -1};
    // NonLinear system: the reverse step at forward step n needs that step's
    // stage arguments: the state u_n (xbase[0]) and a_1..a_3 (xbase[1..3]),
    // recomputed from the segment's stored states (d_seg); xbase[4] takes the
    // stage accumulator, d_w the nonlinear terms' scratch
    bool nl = false;
    double *xbase[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double *d_seg = nullptr; // K dense states of the segment being reversed
    double *d_w = nullptr;   // 4 dense fields
};

namespace {

double *adj_origin(const tsp_adjoint *a, int reg) {
    const tsp_solver *s = a->s;
    return a->abase[reg] + kXOff + (int64_t)s->G * s->sj + (int64_t)s->G * s->sk;
}

double *x_origin(const tsp_adjoint *a, int q) {
    const tsp_solver *s = a->s;
    return a->xbase[q] + kXOff + (int64_t)s->G * s->sj + (int64_t)s->G * s->sk;
}

// the forward step's stage arguments a_1..a_3 from u_n in xbase[0]: forward
// stages 0..2 (the solver's own stage kernels, virtual ghosts) writing their
// next argument into xbase[1..3] (SingleLevelRK4Integrator.hpp:137-156)
tsp_status enqueue_nl_stage_args(tsp_adjoint *a) {
    tsp_solver *s = a->s;
    const double dt = s->d.dt;
    for (int stage = 0; stage < 3; ++stage) {
        tsp::StageArgs g = stage_args(s, stage, 0, 0, 0, 0, dt * kBFrac[stage], dt * kAFrac[stage + 1]);
        g.src = x_origin(a, stage); // u_n, then a_1, a_2
        g.init = x_origin(a, 0);
        g.acc = x_origin(a, 4);
        g.argn = x_origin(a, stage + 1);
        g.ghost_mode = 0; // periodic images through the maps, Dirichlet ghosts written at creation
        hipError_t e = hipErrorNotSupported;
        if (s->launch_stage[stage]) e = s->launch_stage[stage](g, s->cfg.tile_y, s->stream);
        if (e == hipErrorNotSupported) e = s->launch(g, s->cfg.tile_y, s->stream);
        TSP_HIP(e);
    }
    return TSP_OK;
}

int ckpt_slot_of_step(const tsp_adjoint *a, int step) {
    if (step == a->nsteps) return a->nck - 1;
    return step % a->K == 0 ? step / a->K : -1;
}

void drop_adjoint_graphs(tsp_adjoint *a) {
    for (auto *g : {&a->fgraph[0], &a->fgraph[1], &a->bgraph[0], &a->bgraph[1]})
        if (*g) {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

tsp_status adjoint_check_graphs(tsp_adjoint *a) {
    if (std::memcmp(&a->gcfg, &a->s->cfg, sizeof(tsp_kernel_cfg)) != 0) {
        drop_adjoint_graphs(a);
        a->gcfg = a->s->cfg;
    }
    return TSP_OK;
}

tsp_status enqueue_fwd_gather(tsp_adjoint *a) {
    tsp_solver *s = a->s;
    tsp_status st = enqueue_rk4_step(s, nullptr);
    if (st != TSP_OK) return st;
    TSP_HIP(tsp::launch_gather_trace(reg_origin(s, s->role[TSP_ROLE_OLD]), a->d_rec, a->nrec, a->d_trace, a->nsteps,
                                     a->d_counter, s->stream));
    return TSP_OK;
}

// lambda <- G^T lambda on the adjoint registers (lambda in role OLD)
tsp_status enqueue_adjoint_step(tsp_adjoint *a) {
    tsp_solver *s = a->s;
    const double dt = s->d.dt;
    int &OLD = a->arole[TSP_ROLE_OLD], &NEW = a->arole[TSP_ROLE_NEW], &ARG = a->arole[TSP_ROLE_ARG],
        &ARGN = a->arole[TSP_ROLE_ARG_NEXT];
    for (int j = 3; j >= 0; --j) {
        if (j < 3) std::swap(ARG, ARGN);
        tsp::StageArgs g = stage_args(s, j, 0, 0, 0, -1, 0.0, j > 0 ? dt * kAFrac[j] : 0.0);
        g.src = adj_origin(a, j == 3 ? OLD : ARG);
        g.init = adj_origin(a, OLD);
        g.acc = adj_origin(a, NEW);
        g.argn = j > 0 ? adj_origin(a, ARGN) : nullptr;
        g.src_scale = j == 3 ? dt * kBFrac[3] : 1.0;
        g.c_init_argn = j > 0 ? dt * kBFrac[j - 1] : 0.0;
        g.ghost_mode = 0; // adjoint ghosts: periodic images or zero (prefilled)
        for (int c = 0; c < TSP_MAX_COMP; ++c) g.target[c] = 0.0;
        TSP_HIP(a->launch(g, s->cfg.tile_y, s->stream));
        if (a->nl) { // + N^T mu_j at the forward stage argument a_j, in both uses of nu_j
            tsp::NlAdjArgs q{};
            q.mu = g.src;
            q.mu_scale = g.src_scale;
            q.abar = x_origin(a, j);
            q.acc = g.acc;
            q.argn = g.argn;
            q.a_next = g.a_next;
            q.w = a->d_w;
            q.nx = s->d.n[0];
            q.ny = s->d.n[1];
            q.nz = s->d.n[2];
            q.sj = (int32_t)s->sj;
            q.sk = (int32_t)s->sk;
            q.sc = (int32_t)s->sc;
            q.n = s->d.stencil_n;
            for (int d = 0; d < 3; ++d) {
                q.periodic[d] = s->d.periodic[d];
                q.gf1[d] = s->gf1[d];
            }
            q.gamma = s->d.gamma;
            TSP_HIP(tsp::launch_nonlinear_adjoint_terms(q, s->stream));
        }
    }
    std::swap(OLD, NEW);
    return TSP_OK;
}

tsp_status enqueue_inject_adjoint(tsp_adjoint *a) {
    TSP_HIP(tsp::launch_inject_trace(adj_origin(a, a->arole[TSP_ROLE_OLD]), a->d_rec, a->nrec, a->d_dtrace, a->nsteps,
                                     a->d_counter, a->s->stream));
    return enqueue_adjoint_step(a);
}

// Capture `body` twice into a graph for the current parity if it is missing.
template <class F>
tsp_status capture_pair(hipStream_t stream, hipGraphExec_t *slot, int *roles, F &&body) {
    if (*slot) return TSP_OK;
    int saved[4];
    std::memcpy(saved, roles, sizeof(saved));
    TSP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    tsp_status st = body();
    if (st == TSP_OK) st = body();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &graph);
    std::memcpy(roles, saved, sizeof(saved));
    if (st != TSP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(slot, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        *slot = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    return TSP_OK;
}

// forward steps n0 -> n1 with the receiver gather after every step (counter must point at row n0)
tsp_status run_forward_range(tsp_adjoint *a, int n0, int n1) {
    tsp_solver *s = a->s;
    int n = n0;
    while (n < n1) {
        if (n1 - n >= 2) {
            const int parity = (int)((s->steps - a->step0) & 1); // roles are roles0 after an even count
            tsp_status st = capture_pair(s->stream, &a->fgraph[parity], s->role, [&] { return enqueue_fwd_gather(a); });
            if (st != TSP_OK) return st;
            TSP_HIP(hipGraphLaunch(a->fgraph[parity], s->stream));
            advance_clock(s);
            advance_clock(s);
            n += 2;
        } else {
            tsp_status st = enqueue_fwd_gather(a);
            if (st != TSP_OK) return st;
            advance_clock(s);
            n += 1;
        }
    }
    return TSP_OK;
}

// reverse steps n1 -> n0 (counter must point at row n1 - 1)
tsp_status run_backward_range(tsp_adjoint *a, int n1, int n0, int *bparity) {
    int n = n1;
    while (n > n0) {
        if (n - n0 >= 2) {
            tsp_status st = capture_pair(a->s->stream, &a->bgraph[*bparity], a->arole,
                                         [&] { return enqueue_inject_adjoint(a); });
            if (st != TSP_OK) return st;
            TSP_HIP(hipGraphLaunch(a->bgraph[*bparity], a->s->stream));
            n -= 2;
        } else {
            tsp_status st = enqueue_inject_adjoint(a);
            if (st != TSP_OK) return st;
            *bparity ^= 1;
            n -= 1;
        }
    }
    return TSP_OK;
}

// NonLinear system: forward steps n0 -> n1 storing every state u_n (n0 <= n < n1)
// in the segment buffer, with the receiver gather after every step
tsp_status run_forward_range_storing(tsp_adjoint *a, int n0, int n1) {
    tsp_solver *s = a->s;
    for (int n = n0; n < n1; ++n) {
        TSP_HIP(tsp::launch_pack(a->d_seg + (int64_t)(n - n0) * a->state_elems, reg_origin(s, s->role[TSP_ROLE_OLD]),
                                 s->d.n[0], s->d.n[1], s->d.n[2], s->nc, (int)s->sj, (int)s->sk, (int)s->sc,
                                 s->stream));
        tsp_status st = enqueue_fwd_gather(a);
        if (st != TSP_OK) return st;
        advance_clock(s);
    }
    return TSP_OK;
}

// NonLinear system: reverse steps n1 -> n0, each linearised at its stored state
// (counter must point at row n1 - 1)
tsp_status run_backward_range_nl(tsp_adjoint *a, int n1, int n0) {
    tsp_solver *s = a->s;
    for (int n = n1; n > n0; --n) {
        TSP_HIP(tsp::launch_unpack(x_origin(a, 0), a->d_seg + (int64_t)(n - 1 - n0) * a->state_elems, s->d.n[0],
                                   s->d.n[1], s->d.n[2], s->nc, (int)s->sj, (int)s->sk, (int)s->sc, s->stream));
        tsp_status st = enqueue_nl_stage_args(a);
        if (st == TSP_OK) st = enqueue_inject_adjoint(a);
        if (st != TSP_OK) return st;
    }
    return TSP_OK;
}

tsp_status save_checkpoint(tsp_adjoint *a, int slot) {
    tsp_solver *s = a->s;
    TSP_HIP(tsp::launch_pack(a->d_ckpt + (int64_t)slot * a->state_elems, reg_origin(s, s->role[TSP_ROLE_OLD]), s->d.n[0],
                             s->d.n[1], s->d.n[2], s->nc, (int)s->sj, (int)s->sk, (int)s->sc, s->stream));
    return TSP_OK;
}

// put checkpoint `slot` (the state `step` steps into the forward pass) back into the forward solver
tsp_status restore_checkpoint(tsp_adjoint *a, int slot, int step, double time) {
    tsp_solver *s = a->s;
    // the roles `step` RK4 steps after the start of the forward pass: an odd
    // count swaps old <-> new and arg <-> arg_next
    s->steps = a->step0 + step;
    std::memcpy(s->role, a->roles0, sizeof(s->role));
    if (step & 1) {
        std::swap(s->role[TSP_ROLE_OLD], s->role[TSP_ROLE_NEW]);
        std::swap(s->role[TSP_ROLE_ARG], s->role[TSP_ROLE_ARG_NEXT]);
    }
    s->time = time;
    TSP_HIP(tsp::launch_unpack(reg_origin(s, s->role[TSP_ROLE_OLD]), a->d_ckpt + (int64_t)slot * a->state_elems,
                               s->d.n[0], s->d.n[1], s->d.n[2], s->nc, (int)s->sj, (int)s->sk, (int)s->sc, s->stream));
    return TSP_OK;
}

} // namespace

extern "C" {

tsp_status tsp_adjoint_create(tsp_solver *s, const int32_t *receivers, int32_t nrec, int32_t nsteps,
                              int32_t checkpoint_every, tsp_adjoint **out) {
    if (!out) return fail(TSP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (s->d.system != TSP_SYS_WAVE_PHIPI && s->d.system != TSP_SYS_NONLINEAR)
        return fail(TSP_UNSUPPORTED, "the discrete adjoint covers the (phi,Pi) and NonLinear systems");
    if (s->d.system == TSP_SYS_NONLINEAR && (s->d.stencil_n < 1 || s->d.stencil_n > 4))
        return fail(TSP_UNSUPPORTED, "the NonLinear adjoint needs stencil_n in 1..4");
    if (s->d.bc == TSP_BC_CUSTOM)
        return fail(TSP_UNSUPPORTED, "the discrete adjoint transposes the NoOp and Dirichlet boundaries only");
    if (nrec < 1 || !receivers) return fail(TSP_INVALID_ARG, "need at least one receiver");
    if (nsteps < 1) return fail(TSP_INVALID_ARG, "nsteps must be >= 1");
    if (checkpoint_every < 1) return fail(TSP_INVALID_ARG, "checkpoint_every must be >= 1");
    tsp::LaunchFn launch = tsp::lookup_phipi_adjoint(s->d.stencil_n, s->d.ko_order);
    if (!launch) return fail(TSP_UNSUPPORTED, "no adjoint stage kernel for this stencil width / KO order");
    std::vector<int> off(nrec);
    for (int r = 0; r < nrec; ++r) {
        const int i = receivers[3 * r], j = receivers[3 * r + 1], k = receivers[3 * r + 2];
        if (i < 0 || i >= s->d.n[0] || j < 0 || j >= s->d.n[1] || k < 0 || k >= s->d.n[2])
            return fail(TSP_INVALID_ARG, "receiver " + std::to_string(r) + " lies outside the valid box");
        off[r] = (int)(i + j * s->sj + k * s->sk); // component 0 (phi)
    }
    tsp_adjoint *a = new (std::nothrow) tsp_adjoint();
    if (!a) return fail(TSP_OOM, "host allocation failed");
    auto cleanup = [&](tsp_status e) {
        tsp_adjoint_destroy(a);
        return e;
    };
    a->s = s;
    a->nrec = nrec;
    a->nsteps = nsteps;
    a->K = checkpoint_every;
    a->nck = nsteps / checkpoint_every + 1 + (nsteps % checkpoint_every != 0 ? 1 : 0);
    a->state_elems = (int64_t)s->d.n[0] * s->d.n[1] * s->d.n[2] * s->nc;
    a->launch = launch;
    hipError_t e = hipMalloc((void **)&a->d_rec, sizeof(int) * nrec);
    if (e == hipSuccess) e = hipMemcpy(a->d_rec, off.data(), sizeof(int) * nrec, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_trace, sizeof(double) * nrec * (size_t)nsteps);
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_dtrace, sizeof(double) * nrec * (size_t)nsteps);
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_counter, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_ckpt, sizeof(double) * a->state_elems * a->nck);
    for (int r = 0; r < 4 && e == hipSuccess; ++r) {
        e = hipMalloc((void **)&a->abase[r], sizeof(double) * s->reg_elems);
        // adjoint registers: ghosts are zero (homogeneous boundary values) for good
        if (e == hipSuccess) e = hipMemsetAsync(a->abase[r], 0, sizeof(double) * s->reg_elems, s->stream);
    }
    a->nl = s->d.system == TSP_SYS_NONLINEAR;
    if (a->nl) {
        // stage-argument registers: the forward registers' ghosts (zero memory, Dirichlet values)
        for (int q = 0; q < 5 && e == hipSuccess; ++q) {
            e = hipMalloc((void **)&a->xbase[q], sizeof(double) * s->reg_elems);
            if (e == hipSuccess) e = hipMemsetAsync(a->xbase[q], 0, sizeof(double) * s->reg_elems, s->stream);
            if (e == hipSuccess) {
                tsp::GhostArgs g = ghost_args(s, 0);
                g.p = x_origin(a, q);
                e = tsp::launch_ghost_fill(g, s->stream);
            }
        }
        if (e == hipSuccess) e = hipMalloc((void **)&a->d_seg, sizeof(double) * a->state_elems * checkpoint_every);
        if (e == hipSuccess)
            e = hipMalloc((void **)&a->d_w, sizeof(double) * 4 * (size_t)s->d.n[0] * s->d.n[1] * s->d.n[2]);
    }
    if (e == hipSuccess) e = hipMemsetAsync(a->d_trace, 0, sizeof(double) * nrec * (size_t)nsteps, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipErrorOutOfMemory) return cleanup(fail(TSP_OOM, "adjoint buffers do not fit in device memory"));
    if (e != hipSuccess) return cleanup(hip_fail(e, "adjoint allocation"));
    *out = a;
    return TSP_OK;
}

tsp_status tsp_adjoint_destroy(tsp_adjoint *a) {
    if (!a) return TSP_OK;
    if (a->s && a->s->stream) (void)hipStreamSynchronize(a->s->stream);
    drop_adjoint_graphs(a);
    for (void *p : {(void *)a->d_rec, (void *)a->d_trace, (void *)a->d_dtrace, (void *)a->d_counter, (void *)a->d_ckpt,
                    (void *)a->abase[0], (void *)a->abase[1], (void *)a->abase[2], (void *)a->abase[3],
                    (void *)a->xbase[0], (void *)a->xbase[1], (void *)a->xbase[2], (void *)a->xbase[3],
                    (void *)a->xbase[4], (void *)a->d_seg, (void *)a->d_w})
        if (p) (void)hipFree(p);
    delete a;
    return TSP_OK;
}

tsp_status tsp_adjoint_forward(tsp_adjoint *a) {
    if (!a) return fail(TSP_INVALID_ARG, "adjoint is NULL");
    tsp_solver *s = a->s;
    adjoint_check_graphs(a);
    // checkpoints and trace rows count from the state at the start of the pass
    a->step0 = s->steps;
    a->time0 = s->time;
    std::memcpy(a->roles0, s->role, sizeof(a->roles0));
    if (std::memcmp(a->groles0, a->roles0, sizeof(a->roles0)) != 0) { // forward graphs belong to other roles
        for (auto *g : {&a->fgraph[0], &a->fgraph[1]})
            if (*g) {
                (void)hipGraphExecDestroy(*g);
                *g = nullptr;
            }
        std::memcpy(a->groles0, a->roles0, sizeof(a->roles0));
    }
    tsp_status st = save_checkpoint(a, 0);
    if (st != TSP_OK) return st;
    TSP_HIP(tsp::launch_set_counter(a->d_counter, 0, s->stream));
    int n = 0;
    while (n < a->nsteps && st == TSP_OK) {
        const int e = std::min(a->nsteps, (n / a->K + 1) * a->K);
        st = run_forward_range(a, n, e);
        n = e;
        if (st == TSP_OK) {
            const int slot = ckpt_slot_of_step(a, n);
            if (slot >= 0) st = save_checkpoint(a, slot);
        }
    }
    return st;
}

tsp_status tsp_adjoint_trace(tsp_adjoint *a, double *dst, int32_t on_device) {
    if (!a || !dst) return fail(TSP_INVALID_ARG, "adjoint or destination is NULL");
    TSP_HIP(hipStreamSynchronize(a->s->stream));
    TSP_HIP(hipMemcpy(dst, a->d_trace, sizeof(double) * a->nrec * (size_t)a->nsteps,
                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
    return TSP_OK;
}

tsp_status tsp_adjoint_checkpoint(tsp_adjoint *a, int32_t slot, double *dst, int32_t on_device) {
    if (!a || !dst) return fail(TSP_INVALID_ARG, "adjoint or destination is NULL");
    if (slot < 0 || slot >= a->nck) return fail(TSP_INVALID_ARG, "checkpoint slot out of range");
    TSP_HIP(hipStreamSynchronize(a->s->stream));
    TSP_HIP(hipMemcpy(dst, a->d_ckpt + (int64_t)slot * a->state_elems, sizeof(double) * a->state_elems,
                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
    return TSP_OK;
}

tsp_status tsp_adjoint_backward(tsp_adjoint *a, const double *dtrace, int32_t dtrace_on_device, double *grad,
                                int32_t grad_on_device, int32_t recompute) {
    if (!a || !dtrace || !grad) return fail(TSP_INVALID_ARG, "adjoint, dtrace or grad is NULL");
    tsp_solver *s = a->s;
    adjoint_check_graphs(a);
    TSP_HIP(hipMemcpyAsync(a->d_dtrace, dtrace, sizeof(double) * a->nrec * (size_t)a->nsteps,
                           dtrace_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->stream));
    std::memcpy(a->arole, (const int[4]){0, 1, 2, 3}, sizeof(a->arole));
    TSP_HIP(tsp::launch_zero(a->abase[0], s->reg_elems, s->stream)); // lambda = 0 (ghosts stay zero)
    int bparity = 0;
    tsp_status st = TSP_OK;
    if (!recompute && !a->nl) {
        TSP_HIP(tsp::launch_set_counter(a->d_counter, a->nsteps - 1, s->stream));
        st = run_backward_range(a, a->nsteps, 0, &bparity);
    } else {
        // segments from the last to the first: restore the segment's checkpoint,
        // recompute its forward states (and receiver trace), then reverse it
        const double saved_time = s->time;
        const int nseg = (a->nsteps + a->K - 1) / a->K;
        std::vector<double> seg_time(nseg + 1, a->time0);
        for (int m = 1; m <= nseg; ++m) { // the clock at every segment start, accumulated as the steps do
            double t = seg_time[m - 1];
            const int len = std::min(a->K, a->nsteps - (m - 1) * a->K);
            for (int q = 0; q < len; ++q) t += s->d.dt;
            seg_time[m] = t;
        }
        for (int m = nseg - 1; m >= 0 && st == TSP_OK; --m) {
            const int n0 = m * a->K, n1 = std::min(a->nsteps, n0 + a->K);
            st = restore_checkpoint(a, m, n0, seg_time[m]);
            if (st == TSP_OK) {
                hipError_t e = tsp::launch_set_counter(a->d_counter, n0, s->stream);
                if (e != hipSuccess) st = hip_fail(e, "set counter");
            }
            if (st == TSP_OK) st = a->nl ? run_forward_range_storing(a, n0, n1) : run_forward_range(a, n0, n1);
            if (st == TSP_OK) {
                hipError_t e = tsp::launch_set_counter(a->d_counter, n1 - 1, s->stream);
                if (e != hipSuccess) st = hip_fail(e, "set counter");
            }
            if (st == TSP_OK) st = a->nl ? run_backward_range_nl(a, n1, n0) : run_backward_range(a, n1, n0, &bparity);
        }
        // leave the forward solver at the end of the forward pass
        if (st == TSP_OK) st = restore_checkpoint(a, a->nck - 1, a->nsteps, saved_time);
    }
    if (st != TSP_OK) return st;
    const int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    double *dst = grad;
    double *staging = nullptr;
    if (!grad_on_device) {
        TSP_HIP(hipMalloc((void **)&staging, sizeof(double) * a->state_elems));
        dst = staging;
    }
    hipError_t e = tsp::launch_pack(dst, adj_origin(a, a->arole[TSP_ROLE_OLD]), nx, ny, nz, s->nc, (int)s->sj,
                                    (int)s->sk, (int)s->sc, s->stream);
    if (e == hipSuccess && staging)
        e = hipMemcpyAsync(grad, staging, sizeof(double) * a->state_elems, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (staging) (void)hipFree(staging);
    if (e != hipSuccess) return hip_fail(e, "adjoint gradient");
    return TSP_OK;
}

tsp_status tsp_adjoint_apply(tsp_adjoint *a, const double *lam_in, double *lam_out, int32_t on_device) {
    if (!a || !lam_in || !lam_out) return fail(TSP_INVALID_ARG, "adjoint or buffer is NULL");
    tsp_solver *s = a->s;
    const int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    const size_t bytes = sizeof(double) * a->state_elems;
    double *staging = nullptr;
    const double *src = lam_in;
    if (!on_device) {
        TSP_HIP(hipMalloc((void **)&staging, bytes));
        hipError_t e = hipMemcpyAsync(staging, lam_in, bytes, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) {
            (void)hipFree(staging);
            return hip_fail(e, "hipMemcpyAsync");
        }
        src = staging;
    }
    std::memcpy(a->arole, (const int[4]){0, 1, 2, 3}, sizeof(a->arole));
    hipError_t e = tsp::launch_unpack(adj_origin(a, a->arole[TSP_ROLE_OLD]), src, nx, ny, nz, s->nc, (int)s->sj,
                                      (int)s->sk, (int)s->sc, s->stream);
    tsp_status st = e == hipSuccess ? TSP_OK : hip_fail(e, "unpack");
    if (st == TSP_OK && a->nl) { // linearised at the solver's current state
        // the whole register, ghosts included (the same boundary values)
        e = hipMemcpyAsync(a->xbase[0], s->base[s->role[TSP_ROLE_OLD]], sizeof(double) * (size_t)s->reg_elems,
                           hipMemcpyDeviceToDevice, s->stream);
        st = e == hipSuccess ? enqueue_nl_stage_args(a) : hip_fail(e, "state copy");
    }
    if (st == TSP_OK) st = enqueue_adjoint_step(a);
    if (st == TSP_OK) {
        double *dst = on_device ? lam_out : staging;
        e = tsp::launch_pack(dst, adj_origin(a, a->arole[TSP_ROLE_OLD]), nx, ny, nz, s->nc, (int)s->sj, (int)s->sk,
                             (int)s->sc, s->stream);
        if (e == hipSuccess && !on_device) e = hipMemcpyAsync(lam_out, staging, bytes, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) st = hip_fail(e, "adjoint apply");
    }
    if (staging) (void)hipFree(staging);
    return st;
}

} // extern "C"
