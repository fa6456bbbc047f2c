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
#include <new>
#include <string>
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
    double *sigma_tab = nullptr;
    tsp::LaunchFn launch = nullptr;
    tsp_kernel_cfg cfg{0, 8, 64, 1, 1};
    double time = 0.0;
    int64_t steps = 0;
    double gf1[3]{}, gf2[3]{};
    // hipGraph of two RK4 steps per starting parity, valid for `graph_cfg`
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    tsp_kernel_cfg graph_cfg{};
};

namespace {

// Default stage-kernel configuration for a box: tall tiles where the box is
// tall enough, and the longest z chunk (of 64, 32, 16, 8) that still gives at
// least two workgroups per CU, so small boxes fill the device too.
tsp_kernel_cfg default_cfg(const tsp_desc &d) {
    tsp_kernel_cfg c{0, d.n[1] >= 16 ? 16 : 8, 64, 1, 1};
    const long long tiles = (long long)((d.n[0] + 63) / 64) * ((d.n[1] + c.tile_y - 1) / c.tile_y);
    for (int zc : {64, 32, 16, 8}) {
        c.zchunk = std::min(zc, d.n[2]);
        if (tiles * ((d.n[2] + c.zchunk - 1) / c.zchunk) >= 512) break;
    }
    return c;
}

double *reg_origin(const tsp_solver *s, int reg) {
    return s->base[reg] + kXOff + (int64_t)s->G * s->sj + (int64_t)s->G * s->sk;
}

tsp_status validate_cfg(const tsp_kernel_cfg *c) {
    if (!c) return fail(TSP_INVALID_ARG, "kernel config is NULL");
    if (c->ghost_mode != 0 && c->ghost_mode != 1) return fail(TSP_INVALID_ARG, "ghost_mode must be 0 or 1");
    if (c->tile_y != 2 && c->tile_y != 4 && c->tile_y != 8 && c->tile_y != 16)
        return fail(TSP_INVALID_ARG, "tile_y must be 2, 4, 8 or 16");
    if (c->zchunk < 1) return fail(TSP_INVALID_ARG, "zchunk must be >= 1");
    if (c->xcd_remap != 0 && c->xcd_remap != 1) return fail(TSP_INVALID_ARG, "xcd_remap must be 0 or 1");
    if (c->zalt != 0 && c->zalt != 1) return fail(TSP_INVALID_ARG, "zalt must be 0 or 1");
    return TSP_OK;
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
    a.zchunk = s->cfg.zchunk;
    a.xcd_remap = s->cfg.xcd_remap;
    a.zalt = s->cfg.zalt;
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
    const double dt = s->d.dt;
    int &OLD = s->role[TSP_ROLE_OLD], &NEW = s->role[TSP_ROLE_NEW], &ARG = s->role[TSP_ROLE_ARG],
        &ARGN = s->role[TSP_ROLE_ARG_NEXT];
    auto fill = [&](int reg, int slot) -> tsp_status {
        if (s->cfg.ghost_mode != 1) return TSP_OK;
        if (tm) TSP_HIP(hipEventRecord(tm->fill[slot].start, s->stream));
        TSP_HIP(tsp::launch_ghost_fill(ghost_args(s, reg), s->stream));
        if (tm) TSP_HIP(hipEventRecord(tm->fill[slot].stop, s->stream));
        return TSP_OK;
    };
    tsp_status st = fill(OLD, 0); // pre_step
    if (st != TSP_OK) return st;
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
        TSP_HIP(s->launch(a, s->cfg.tile_y, s->stream));
        if (tm) TSP_HIP(hipEventRecord(tm->stage[stage].stop, s->stream));
    }
    std::swap(OLD, NEW); // post_step
    return TSP_OK;
}

tsp_status enqueue_euler_step(tsp_solver *s) {
    int &OLD = s->role[TSP_ROLE_OLD], &NEW = s->role[TSP_ROLE_NEW];
    if (s->cfg.ghost_mode == 1) TSP_HIP(tsp::launch_ghost_fill(ghost_args(s, OLD), s->stream));
    const tsp::StageArgs a = stage_args(s, tsp::kEuler, OLD, OLD, NEW, -1, s->d.dt, 0.0);
    TSP_HIP(s->launch(a, s->cfg.tile_y, s->stream));
    std::swap(OLD, NEW);
    return TSP_OK;
}

void advance_clock(tsp_solver *s) { // Integrator.hpp:98-100
    s->steps++;
    s->time += s->d.dt;
}

} // namespace

extern "C" {

const char *tsp_last_error(void) { return g_err.c_str(); }
const char *tsp_version(void) { return TSP_VERSION; }

int64_t tsp_rk4_bytes_per_cell(const tsp_desc *d) {
    const int nc = d ? system_ncomp(d->system) : -1;
    return nc < 0 ? -1 : 128LL * nc;
}

tsp_status tsp_desc_default(tsp_desc *d, int32_t system, int32_t n) {
    if (!d) return fail(TSP_INVALID_ARG, "descriptor is NULL");
    if (system_ncomp(system) < 0) return fail(TSP_INVALID_ARG, "unknown system id");
    std::memset(d, 0, sizeof(*d));
    d->system = system;
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
    tsp_solver *s = new (std::nothrow) tsp_solver();
    if (!s) return fail(TSP_OOM, "host allocation failed");
    s->d = *d;
    s->nc = system_ncomp(d->system);
    s->G = d->ghost;
    s->H = std::max(d->stencil_n, d->ko_order);
    s->pitch = (int)(((int64_t)kXOff + d->n[0] + s->G + 15) / 16 * 16);
    s->sj = s->pitch;
    s->sk = s->sj * (d->n[1] + 2 * s->G);
    s->sc = s->sk * (d->n[2] + 2 * s->G);
    s->reg_elems = s->sc * s->nc;
    s->launch = lookup_launcher(d->system, d->stencil_n, d->ko_order);
    s->cfg = default_cfg(*d);
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

tsp_status tsp_fill_ghosts(tsp_solver *s, int32_t role) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (role < 0 || role > 3) return fail(TSP_INVALID_ARG, "role must be 0..3");
    TSP_HIP(tsp::launch_ghost_fill(ghost_args(s, s->role[role]), s->stream));
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
    s->cfg = *cfg;
    return TSP_OK;
}

tsp_status tsp_sync(tsp_solver *s) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    TSP_HIP(hipStreamSynchronize(s->stream));
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
    if (std::memcmp(&s->graph_cfg, &s->cfg, sizeof(tsp_kernel_cfg)) != 0) {
        drop_graphs(s);
        s->graph_cfg = s->cfg;
    }
    const int parity = (int)(s->steps & 1);
    if (!s->graph[parity] && nsteps >= 2) {
        int saved[4];
        std::memcpy(saved, s->role, sizeof(saved));
        TSP_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        tsp_status st = enqueue_rk4_step(s, nullptr);
        if (st == TSP_OK) st = enqueue_rk4_step(s, nullptr);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &graph);
        std::memcpy(s->role, saved, sizeof(saved)); // two steps bring the roles back
        if (st != TSP_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&s->graph[parity], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            s->graph[parity] = nullptr;
            return hip_fail(e, "hipGraphInstantiate");
        }
    }
    int64_t n = 0;
    for (; n + 2 <= nsteps; n += 2) {
        TSP_HIP(hipGraphLaunch(s->graph[parity], s->stream));
        advance_clock(s);
        advance_clock(s);
    }
    if (n < nsteps) return tsp_rk4_steps(s, nsteps - n);
    return TSP_OK;
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
            if (s->cfg.ghost_mode == 1 && fill_ms && st == TSP_OK) {
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

bool cache_lookup(const std::string &key, tsp_kernel_cfg *cfg) {
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
        if (std::sscanf(line, "%255s %d %d %d %d %d %lf", k, &c.ghost_mode, &c.tile_y, &c.zchunk, &c.xcd_remap, &c.zalt,
                        &ms) == 7 &&
            key == k && validate_cfg(&c) == TSP_OK) {
            *cfg = c; // the last matching line wins
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
    std::fprintf(f, "%s %d %d %d %d %d %.6f\n", key.c_str(), c.ghost_mode, c.tile_y, c.zchunk, c.xcd_remap, c.zalt, ms);
    std::fclose(f);
}

} // namespace

tsp_status tsp_autotune(tsp_solver *s, int32_t steps, tsp_kernel_cfg *chosen) {
    if (!s) return fail(TSP_INVALID_ARG, "solver is NULL");
    if (steps < 1) steps = 1;
    const std::string key = tune_key(s);
    tsp_kernel_cfg best = s->cfg;
    if (cache_lookup(key, &best)) {
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
    // candidates: every tile height, z chunk, XCD mapping and z-march order
    // with virtual ghosts, then the best geometry with the fill-pass mode
    std::vector<tsp_kernel_cfg> cands;
    for (int ty : {2, 4, 8, 16})
        for (int zc : zchunks)
            for (int xr = 0; xr < 2; ++xr)
                for (int za = 0; za < 2; ++za) cands.push_back(tsp_kernel_cfg{0, ty, zc, xr, za});
    double best_ms = 1e300;
    auto time_cfg = [&](const tsp_kernel_cfg &c) -> tsp_status {
        s->cfg = c;
        tsp_status r = enqueue_rk4_step(s, nullptr); // warm-up
        if (r != TSP_OK) return r;
        if ((e = hipEventRecord(e0, s->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        for (int n = 0; n < steps; ++n)
            if ((r = enqueue_rk4_step(s, nullptr)) != TSP_OK) return r;
        if ((e = hipEventRecord(e1, s->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((e = hipEventSynchronize(e1)) != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
        const double per = ms / steps;
        if (per < best_ms) {
            best_ms = per;
            best = c;
        }
        // restart every candidate from the saved state so values stay bounded
        std::memcpy(s->role, saved_role, sizeof(saved_role));
        e = hipMemcpyAsync(s->base[saved_role[TSP_ROLE_OLD]], backup, sizeof(double) * s->reg_elems,
                           hipMemcpyDeviceToDevice, s->stream);
        return e == hipSuccess ? TSP_OK : hip_fail(e, "autotune restore");
    };
    for (size_t i = 0; i < cands.size() && st == TSP_OK; ++i) st = time_cfg(cands[i]);
    if (st == TSP_OK) {
        tsp_kernel_cfg fill = best;
        fill.ghost_mode = 1;
        st = time_cfg(fill);
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
