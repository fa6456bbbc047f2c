// batch.hip -- the batched half of the C ABI (tsp_batch_*) and the batched
// stage kernels of the (phi, Pi) system (generic body).
//
// A batch is many independent single-box models advanced together. All
// registers of all instances live in one pool allocation (each register keeps
// the single-box layout, see capi.hip); a device instance table and a work
// list of (instance, tile, z chunk) items drive one persistent launch per RK4
// stage. The reference has no batched mode: this is the data-parallel way to
// run an ensemble of SingleLevelRK4Integrator solves
// (integrators/SingleLevelRK4Integrator.hpp:95-157) of uneven sizes.
#include "batch_kernel.hpp"
#include "capi_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

namespace tsp {
// forward stages 0 and 3 run their stage-specific bodies
hipError_t launch_batched_phipi(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                                hipStream_t stream) {
    if (b.stage == kStage0) return launch_batched_phipi_s0(variant, b, tile_y, grid, max_abl_T, halo_ring, stream);
    if (b.stage == kStage3) return launch_batched_phipi_s3(variant, b, tile_y, grid, max_abl_T, halo_ring, stream);
    return launch_batched_phipi_mode<-1>(variant, b, tile_y, grid, max_abl_T, halo_ring, stream);
}
hipError_t launch_batched_group(int group, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                                hipStream_t stream) {
    const int v = group % kBatchVariants;
    switch (group / kBatchVariants) {
    case 0: return launch_batched_phipi(v, b, tile_y, grid, max_abl_T, halo_ring, stream);
    case 1: return launch_batched_phiu(v, b, tile_y, grid, max_abl_T, halo_ring, stream);
    case 2: return launch_batched_nonlinear(v, b, tile_y, grid, max_abl_T, halo_ring, stream);
    default: return hipErrorInvalidValue;
    }
}
} // namespace tsp

using namespace tsp_host;

// one kernel configuration per work-list group (system, stencil width, KO order)
struct GroupCfgs {
    tsp_batch_cfg g[tsp::kBatchGroups];
    bool operator==(const GroupCfgs &o) const { return std::memcmp(g, o.g, sizeof(g)) == 0; }
    bool operator!=(const GroupCfgs &o) const { return !(*this == o); }
    void fill(const tsp_batch_cfg &c) {
        for (auto &x : g) x = c;
    }
};

struct tsp_batch {
    int count = 0;
    std::vector<tsp_desc> d;
    std::vector<Layout> L;
    std::vector<int64_t> reg_off; // 4 per instance: element offset of each register's base in the pool
    double *pool = nullptr;
    int64_t pool_elems = 0;
    std::vector<double *> sigma;  // per instance (null without ABL)
    int max_abl_T = 0;            // thickest absorbing layer (sizes the kernels' LDS rate table)
    std::vector<tsp::BatchInstance> h_inst;
    tsp::BatchInstance *d_inst = nullptr;
    tsp::BatchWork *d_work = nullptr;
    int nwork = 0;
    int group_off[tsp::kBatchGroups] = {}; // work items of group v: [group_off[v], group_off[v] + group_n[v])
    int group_n[tsp::kBatchGroups] = {};
    int role[4] = {0, 1, 2, 3};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    tsp_batch_cfg cfg{16, 0, 0, 1, 1, 1}; // what tsp_batch_get_kernel reports (the largest group's)
    GroupCfgs gcfg;                        // the kernel configuration of every work-list group
    GroupCfgs work_gcfg;                   // configurations the work list was built for
    bool work_valid = false;
    int only_group = -1;                   // >= 0: steps launch this group only (the autotuner)
    std::vector<double> time;
    int64_t steps = 0;
    // two-step graphs, each for the register roles it was captured with
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_role[2][4] = {};
    int graph_next = 0;
    GroupCfgs graph_gcfg;                  // configurations the graphs were captured with
    int graph_only = -2;                   // only_group the graphs were captured with (-2: none captured)
    uint64_t work_gen = 0; // bumped whenever the work list is rebuilt (graphs of other objects hold it)
    // a level (tsp_level_create): the instances are the boxes of one domain and
    // every stage starts with the ghost exchange between them
    bool level = false;
    tsp_desc ldesc{};            // the domain
    tsp::LevelArgs largs{};      // device view of the box layout
    int32_t *d_axis_box = nullptr;
    long long max_shell = 0;     // ghost cells of the largest box
    std::vector<int32_t> box_lo; // 3 per box
};

namespace {

double *batch_origin(const tsp_batch *b, int i, int reg) {
    const Layout &L = b->L[i];
    return b->pool + b->reg_off[4 * i + reg] + kXOff + (int64_t)L.G * L.sj + (int64_t)L.G * L.sk;
}

tsp_status validate_batch_cfg(const tsp_batch_cfg *c) {
    if (!c) return fail(TSP_INVALID_ARG, "batch config is NULL");
    if (c->tile_y != 8 && c->tile_y != 16) return fail(TSP_INVALID_ARG, "batch tile_y must be 8 or 16");
    if (c->zchunk < 0) return fail(TSP_INVALID_ARG, "zchunk must be >= 0 (0: automatic)");
    if (c->grid_blocks < 0) return fail(TSP_INVALID_ARG, "grid_blocks must be >= 0 (0: one block per work item)");
    if (c->zalt < 0 || c->zalt > 2) return fail(TSP_INVALID_ARG, "zalt must be 0, 1 or 2");
    if (c->halo_ring != 0 && c->halo_ring != 1) return fail(TSP_INVALID_ARG, "halo_ring must be 0 or 1");
    if (c->narrow_tiles != 0 && c->narrow_tiles != 1) return fail(TSP_INVALID_ARG, "narrow_tiles must be 0 or 1");
    return TSP_OK;
}

int auto_zchunk(int nz) { return std::min(nz, 32); }

constexpr int kIdentityRoles[4] = {0, 1, 2, 3};
// register roles one RK4 step after `r` (old <-> new, arg <-> arg_next)
void rk4_next_roles(const int r[4], int out[4]) {
    out[TSP_ROLE_OLD] = r[TSP_ROLE_NEW];
    out[TSP_ROLE_NEW] = r[TSP_ROLE_OLD];
    out[TSP_ROLE_ARG] = r[TSP_ROLE_ARG_NEXT];
    out[TSP_ROLE_ARG_NEXT] = r[TSP_ROLE_ARG];
}

void drop_batch_graphs(tsp_batch *b) {
    for (auto &g : b->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    b->graph_only = -2;
    b->graph_next = 0;
}

// (Re)build the device work list: per group, instances in decreasing cell
// count, each split into 64 x tile_y column tiles and z chunks of the group's
// configuration; with narrow_tiles, an x remainder of at most 32 cells becomes
// a strip of 32 x 2*tile_y tiles.
tsp_status build_work(tsp_batch *b) {
    if (b->d_work && b->work_valid && b->work_gcfg == b->gcfg) return TSP_OK;
    std::vector<int> order(b->count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int64_t cx = (int64_t)b->d[x].n[0] * b->d[x].n[1] * b->d[x].n[2];
        const int64_t cy = (int64_t)b->d[y].n[0] * b->d[y].n[1] * b->d[y].n[2];
        return cx > cy;
    });
    std::vector<tsp::BatchWork> work;
    for (int v = 0; v < tsp::kBatchGroups; ++v) {
        const tsp_batch_cfg &gc = b->gcfg.g[v];
        const int ty = gc.tile_y;
        b->group_off[v] = (int)work.size();
        for (int i : order) {
            if (b->h_inst[i].variant != v) continue;
            const tsp_desc &d = b->d[i];
            const int zc = gc.zchunk > 0 ? std::min(gc.zchunk, d.n[2]) : auto_zchunk(d.n[2]);
            b->h_inst[i].zchunk = zc;
            const int nzc = (d.n[2] + zc - 1) / zc;
            // x origin of the narrow strip (nx if there is none)
            const int xn = gc.narrow_tiles && d.n[0] % 64 != 0 && d.n[0] % 64 <= 32 ? d.n[0] / 64 * 64 : d.n[0];
            for (int c = 0; c < nzc; ++c) {
                for (int y0 = 0; y0 < d.n[1]; y0 += ty)
                    for (int x0 = 0; x0 < xn; x0 += 64) work.push_back(tsp::BatchWork{i, x0, y0, c, 0});
                if (xn < d.n[0])
                    for (int y0 = 0; y0 < d.n[1]; y0 += 2 * ty) work.push_back(tsp::BatchWork{i, xn, y0, c, 1});
            }
        }
        b->group_n[v] = (int)work.size() - b->group_off[v];
    }
    if (work.size() > 0x7fffffff) return fail(TSP_UNSUPPORTED, "work list too long");
    if (b->d_work) {
        // captured graphs hold pointers into the old work list
        TSP_HIP(hipStreamSynchronize(b->stream));
        drop_batch_graphs(b);
        (void)hipFree(b->d_work);
        b->d_work = nullptr;
    }
    TSP_HIP(hipMalloc((void **)&b->d_work, sizeof(tsp::BatchWork) * work.size()));
    TSP_HIP(hipMemcpy(b->d_work, work.data(), sizeof(tsp::BatchWork) * work.size(), hipMemcpyHostToDevice));
    TSP_HIP(hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(tsp::BatchInstance) * b->count, hipMemcpyHostToDevice));
    b->nwork = (int)work.size();
    b->work_gcfg = b->gcfg;
    b->work_valid = true;
    b->work_gen++;
    return TSP_OK;
}

// One RK4 step of every instance: per stage, one persistent launch per
// (system, stencil width, KO order) group present in the batch.
tsp_status enqueue_batch_step(tsp_batch *b) {
    int &OLD = b->role[TSP_ROLE_OLD], &NEW = b->role[TSP_ROLE_NEW], &ARG = b->role[TSP_ROLE_ARG],
        &ARGN = b->role[TSP_ROLE_ARG_NEXT];
    for (int stage = 0; stage < 4; ++stage) {
        if (stage > 0) std::swap(ARG, ARGN);
        if (b->level) // FillPatch of the stage's source register across the boxes
            TSP_HIP(tsp::launch_level_exchange(b->largs, stage == 0 ? OLD : ARG, b->max_shell, b->stream));
        for (int v = 0; v < tsp::kBatchGroups; ++v) {
            if (b->group_n[v] == 0 || (b->only_group >= 0 && v != b->only_group)) continue;
            const tsp_batch_cfg &gc = b->gcfg.g[v];
            tsp::BatchArgs a{};
            a.inst = b->d_inst;
            a.work = b->d_work + b->group_off[v];
            a.nwork = b->group_n[v];
            a.stage = stage;
            a.src = stage == 0 ? OLD : ARG;
            a.init = OLD;
            a.acc = NEW;
            a.argn = stage < 3 ? ARGN : -1;
            a.zalt = gc.zalt;
            a.ghost_mode = b->level ? 1 : 0; // a box's ghosts hold its neighbours' values after the exchange
            const int grid = gc.grid_blocks > 0 ? std::min(gc.grid_blocks, a.nwork) : a.nwork;
            TSP_HIP(tsp::launch_batched_group(v, a, gc.tile_y, grid, b->max_abl_T, gc.halo_ring, b->stream));
        }
    }
    std::swap(OLD, NEW);
    return TSP_OK;
}

// One forward Euler step of every instance (SingleLevelEulerIntegrator.hpp:66-88).
tsp_status enqueue_batch_euler_step(tsp_batch *b) {
    int &OLD = b->role[TSP_ROLE_OLD], &NEW = b->role[TSP_ROLE_NEW];
    if (b->level) TSP_HIP(tsp::launch_level_exchange(b->largs, OLD, b->max_shell, b->stream));
    for (int v = 0; v < tsp::kBatchGroups; ++v) {
        if (b->group_n[v] == 0 || (b->only_group >= 0 && v != b->only_group)) continue;
        const tsp_batch_cfg &gc = b->gcfg.g[v];
        tsp::BatchArgs a{};
        a.inst = b->d_inst;
        a.work = b->d_work + b->group_off[v];
        a.nwork = b->group_n[v];
        a.stage = tsp::kEuler;
        a.src = OLD;
        a.init = OLD;
        a.acc = NEW;
        a.argn = -1;
        a.zalt = gc.zalt;
        a.ghost_mode = b->level ? 1 : 0;
        const int grid = gc.grid_blocks > 0 ? std::min(gc.grid_blocks, a.nwork) : a.nwork;
        TSP_HIP(tsp::launch_batched_group(v, a, gc.tile_y, grid, b->max_abl_T, gc.halo_ring, b->stream));
    }
    std::swap(OLD, NEW);
    return TSP_OK;
}

void batch_advance_clock(tsp_batch *b) {
    b->steps++;
    for (int i = 0; i < b->count; ++i) b->time[i] += b->d[i].dt;
}


// The captured two-step graph for register roles `roles` (captured with those
// roles if missing; two RK4 steps bring the roles back); graphs of another
// configuration are dropped first.
tsp_status ensure_batch_graph(tsp_batch *b, const int roles[4], int *slot) {
    if (b->graph_only != b->only_group || b->graph_gcfg != b->gcfg) {
        drop_batch_graphs(b);
        b->graph_gcfg = b->gcfg;
        b->graph_only = b->only_group;
    }
    for (int g = 0; g < 2; ++g)
        if (b->graph[g] && std::memcmp(b->graph_role[g], roles, sizeof(b->graph_role[g])) == 0) {
            *slot = g;
            return TSP_OK;
        }
    const int g = !b->graph[0] ? 0 : !b->graph[1] ? 1 : b->graph_next;
    b->graph_next = 1 - g;
    if (b->graph[g]) {
        (void)hipGraphExecDestroy(b->graph[g]);
        b->graph[g] = nullptr;
    }
    int saved[4];
    std::memcpy(saved, b->role, sizeof(saved));
    std::memcpy(b->role, roles, sizeof(saved));
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        std::memcpy(b->role, saved, sizeof(saved));
        return hip_fail(e, "hipStreamBeginCapture");
    }
    tsp_status st = enqueue_batch_step(b);
    if (st == TSP_OK) st = enqueue_batch_step(b);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(b->stream, &graph);
    std::memcpy(b->role, saved, sizeof(saved));
    if (st != TSP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->graph[g], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        b->graph[g] = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    std::memcpy(b->graph_role[g], roles, sizeof(b->graph_role[g]));
    *slot = g;
    return TSP_OK;
}
} // namespace

extern "C" {

namespace {
// A batch of `count` instances. `box_off`/`dom` (3 per instance / 3): the
// instances are the boxes of one validated domain (a level): per-instance
// descriptors are then only checked for what the batched kernels need.
tsp_status create_batch(const tsp_desc *descs, int32_t count, void *stream, tsp_batch **out, const int32_t *box_off,
                        const int32_t *dom) {
    if (!out) return fail(TSP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!descs || count < 1) return fail(TSP_INVALID_ARG, "need at least one instance");
    for (int i = 0; i < count; ++i) {
        tsp_status st = box_off ? TSP_OK : validate(&descs[i]);
        if (st != TSP_OK) return fail(st, "instance " + std::to_string(i) + ": " + g_err);
        if (descs[i].bc == TSP_BC_CUSTOM)
            return fail(TSP_UNSUPPORTED, "instance " + std::to_string(i) +
                                             ": batches and levels carry NoOp and Dirichlet boundaries only");
        if (tsp::batch_system_slot(descs[i].system) < 0)
            return fail(TSP_UNSUPPORTED, "instance " + std::to_string(i) +
                                             ": batched mode runs the (phi,Pi), (phi,u) and NonLinear systems");
        if (tsp::batch_variant(descs[i].stencil_n, descs[i].ko_order) < 0)
            return fail(TSP_UNSUPPORTED, "instance " + std::to_string(i) + ": no batched kernel for stencil_n=" +
                                             std::to_string(descs[i].stencil_n) +
                                             " ko_order=" + std::to_string(descs[i].ko_order));
        // the batched kernels address registers through 4 GiB buffer descriptors
        const Layout L = layout_of(descs[i]);
        if ((L.elems + (int64_t)L.G * (1 + L.sj + L.sk)) * (int64_t)sizeof(double) >= (int64_t(1) << 32))
            return fail(TSP_UNSUPPORTED, "instance " + std::to_string(i) + ": batched instances need registers under 4 GiB");
    }
    {
        int maxT = 0;
        for (int i = 0; i < count; ++i) maxT = std::max(maxT, descs[i].abl_thickness);
        tsp_status st = check_abl_lds(maxT);
        if (st != TSP_OK) return st;
    }
    tsp_batch *b = new (std::nothrow) tsp_batch();
    if (b) b->gcfg.fill(b->cfg);
    if (!b) return fail(TSP_OOM, "host allocation failed");
    auto cleanup = [&](tsp_status e) {
        tsp_batch_destroy(b);
        return e;
    };
    b->count = count;
    b->d.assign(descs, descs + count);
    b->time.assign(count, 0.0);
    b->sigma.assign(count, nullptr);
    int64_t off = 0;
    for (int i = 0; i < count; ++i) {
        b->L.push_back(layout_of(descs[i]));
        for (int r = 0; r < 4; ++r) {
            b->reg_off.push_back(off);
            off += (b->L[i].elems + 31) / 32 * 32; // 256-byte aligned registers
        }
    }
    b->pool_elems = off;
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipStreamCreateWithFlags"));
        b->own_stream = true;
    }
    hipError_t e = hipMalloc((void **)&b->pool, sizeof(double) * b->pool_elems);
    if (e == hipErrorOutOfMemory) return cleanup(fail(TSP_OOM, "hipMalloc of the register pool failed (out of memory)"));
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(pool)"));
    e = hipMemsetAsync(b->pool, 0, sizeof(double) * b->pool_elems, b->stream);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMemsetAsync(pool)"));
    b->h_inst.resize(count);
    for (int i = 0; i < count; ++i) {
        const tsp_desc &d = descs[i];
        const Layout &L = b->L[i];
        if (d.abl_thickness > 0) {
            std::vector<double> tab = abl_sigma_table(d);
            e = hipMalloc((void **)&b->sigma[i], sizeof(double) * tab.size());
            if (e == hipSuccess)
                e = hipMemcpy(b->sigma[i], tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) return cleanup(hip_fail(e, "sigma table"));
        }
        tsp::BatchInstance &I = b->h_inst[i];
        std::memset(&I, 0, sizeof(I));
        for (int r = 0; r < 4; ++r) I.reg[r] = batch_origin(b, i, r);
        I.sigma_tab = b->sigma[i];
        I.nx = d.n[0];
        I.ny = d.n[1];
        I.nz = d.n[2];
        I.G = L.G;
        I.sj = (int32_t)L.sj;
        I.sk = (int32_t)L.sk;
        I.sc = (int32_t)L.sc;
        I.variant = tsp::batch_group(d.system, d.stencil_n, d.ko_order);
        for (int a = 0; a < 3; ++a) I.periodic[a] = d.periodic[a];
        I.dirichlet = d.bc == TSP_BC_DIRICHLET;
        I.abl_T = d.abl_thickness;
        for (int a = 0; a < 3; ++a) {
            I.off[a] = box_off ? box_off[3 * i + a] : 0;
            I.dom[a] = dom ? dom[a] : d.n[a];
        }
        b->max_abl_T = std::max(b->max_abl_T, d.abl_thickness);
        I.dt = d.dt;
        for (int a = 0; a < 3; ++a) {
            // Geometry::CellSize of the domain (a box of a level keeps the domain's prob_lo/hi)
            const double dx = (d.prob_hi[a] - d.prob_lo[a]) / (double)(dom ? dom[a] : d.n[a]);
            I.gf1[a] = 1.0 / dx;
            I.gf2[a] = 1.0 / (dx * dx);
        }
        I.ko_strength = d.ko_strength;
        I.gamma = d.gamma;
        for (int c = 0; c < TSP_MAX_COMP; ++c) I.target[c] = d.target[c];
        // physical ghosts hold their boundary values for good (see tsp_create)
        for (int r = 0; r < 4; ++r) {
            tsp::GhostArgs g{};
            g.p = I.reg[r];
            g.nx = d.n[0];
            g.ny = d.n[1];
            g.nz = d.n[2];
            g.G = L.G;
            g.nc = system_ncomp(d.system);
            g.sj = I.sj;
            g.sk = I.sk;
            g.sc = I.sc;
            for (int a = 0; a < 3; ++a) g.periodic[a] = d.periodic[a];
            g.dirichlet = I.dirichlet;
            for (int c = 0; c < TSP_MAX_COMP; ++c) g.target[c] = d.target[c];
            e = tsp::launch_ghost_fill(g, b->stream);
            if (e != hipSuccess) return cleanup(hip_fail(e, "ghost prefill"));
        }
    }
    e = hipMalloc((void **)&b->d_inst, sizeof(tsp::BatchInstance) * count);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(instances)"));
    tsp_status st = build_work(b);
    if (st != TSP_OK) return cleanup(st);
    e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipStreamSynchronize"));
    *out = b;
    return TSP_OK;
}
} // namespace

tsp_status tsp_batch_create(const tsp_desc *descs, int32_t count, void *stream, tsp_batch **out) {
    return create_batch(descs, count, stream, out, nullptr, nullptr);
}

tsp_status tsp_batch_destroy(tsp_batch *b) {
    if (!b) return TSP_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_batch_graphs(b);
    if (b->pool) (void)hipFree(b->pool);
    for (double *p : b->sigma)
        if (p) (void)hipFree(p);
    if (b->d_inst) (void)hipFree(b->d_inst);
    if (b->d_work) (void)hipFree(b->d_work);
    if (b->d_axis_box) (void)hipFree(b->d_axis_box);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return TSP_OK;
}

tsp_status tsp_batch_size(tsp_batch *b, int32_t *count, int64_t *cells) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    if (count) *count = b->count;
    if (cells) {
        int64_t c = 0;
        for (const tsp_desc &d : b->d) c += (int64_t)d.n[0] * d.n[1] * d.n[2];
        *cells = c;
    }
    return TSP_OK;
}

tsp_status tsp_batch_set_state(tsp_batch *b, int32_t index, const double *src, int32_t on_device) {
    if (!b || !src) return fail(TSP_INVALID_ARG, "batch or source is NULL");
    if (index < 0 || index >= b->count) return fail(TSP_INVALID_ARG, "instance index out of range");
    const tsp_desc &d = b->d[index];
    const Layout &L = b->L[index];
    const int nc = system_ncomp(d.system);
    const size_t bytes = sizeof(double) * (size_t)d.n[0] * d.n[1] * d.n[2] * nc;
    const double *dsrc = src;
    double *staging = nullptr;
    if (!on_device) {
        TSP_HIP(hipMalloc((void **)&staging, bytes));
        hipError_t e = hipMemcpyAsync(staging, src, bytes, hipMemcpyHostToDevice, b->stream);
        if (e != hipSuccess) {
            (void)hipFree(staging);
            return hip_fail(e, "hipMemcpyAsync(state)");
        }
        dsrc = staging;
    }
    hipError_t e = tsp::launch_unpack(batch_origin(b, index, b->role[TSP_ROLE_OLD]), dsrc, d.n[0], d.n[1], d.n[2], nc,
                                      (int)L.sj, (int)L.sk, (int)L.sc, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (staging) (void)hipFree(staging);
    if (e != hipSuccess) return hip_fail(e, "batch set_state");
    return TSP_OK;
}

tsp_status tsp_batch_get_state(tsp_batch *b, int32_t index, double *dst, int32_t on_device) {
    if (!b || !dst) return fail(TSP_INVALID_ARG, "batch or destination is NULL");
    if (index < 0 || index >= b->count) return fail(TSP_INVALID_ARG, "instance index out of range");
    const tsp_desc &d = b->d[index];
    const Layout &L = b->L[index];
    const int nc = system_ncomp(d.system);
    const size_t bytes = sizeof(double) * (size_t)d.n[0] * d.n[1] * d.n[2] * nc;
    double *ddst = dst;
    double *staging = nullptr;
    if (!on_device) {
        TSP_HIP(hipMalloc((void **)&staging, bytes));
        ddst = staging;
    }
    hipError_t e = tsp::launch_pack(ddst, batch_origin(b, index, b->role[TSP_ROLE_OLD]), d.n[0], d.n[1], d.n[2], nc,
                                    (int)L.sj, (int)L.sk, (int)L.sc, b->stream);
    if (e == hipSuccess && staging) e = hipMemcpyAsync(dst, staging, bytes, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (staging) (void)hipFree(staging);
    if (e != hipSuccess) return hip_fail(e, "batch get_state");
    return TSP_OK;
}

tsp_status tsp_batch_get_time(tsp_batch *b, int32_t index, double *time, int64_t *steps) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    if (index < 0 || index >= b->count) return fail(TSP_INVALID_ARG, "instance index out of range");
    if (time) *time = b->time[index];
    if (steps) *steps = b->steps;
    return TSP_OK;
}

tsp_status tsp_batch_get_kernel(tsp_batch *b, tsp_batch_cfg *cfg) {
    if (!b || !cfg) return fail(TSP_INVALID_ARG, "batch or cfg is NULL");
    *cfg = b->cfg;
    return TSP_OK;
}

tsp_status tsp_batch_set_kernel(tsp_batch *b, const tsp_batch_cfg *cfg) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    tsp_status st = validate_batch_cfg(cfg);
    if (st != TSP_OK) return st;
    b->cfg = *cfg;
    b->gcfg.fill(*cfg);
    return build_work(b);
}

namespace {
tsp_status group_of(tsp_batch *b, int32_t system, int32_t stencil_n, int32_t ko_order, int *v) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    *v = tsp::batch_group(system, stencil_n, ko_order);
    if (*v < 0)
        return fail(TSP_INVALID_ARG, "no batched kernels for system " + std::to_string(system) + ", stencil_n " +
                                         std::to_string(stencil_n) + ", ko_order " + std::to_string(ko_order));
    return TSP_OK;
}
} // namespace

tsp_status tsp_batch_get_group_kernel(tsp_batch *b, int32_t system, int32_t stencil_n, int32_t ko_order,
                                      tsp_batch_cfg *cfg) {
    int v = -1;
    tsp_status st = group_of(b, system, stencil_n, ko_order, &v);
    if (st != TSP_OK) return st;
    if (!cfg) return fail(TSP_INVALID_ARG, "cfg is NULL");
    *cfg = b->gcfg.g[v];
    return TSP_OK;
}

tsp_status tsp_batch_set_group_kernel(tsp_batch *b, int32_t system, int32_t stencil_n, int32_t ko_order,
                                      const tsp_batch_cfg *cfg) {
    int v = -1;
    tsp_status st = group_of(b, system, stencil_n, ko_order, &v);
    if (st != TSP_OK) return st;
    if ((st = validate_batch_cfg(cfg)) != TSP_OK) return st;
    b->gcfg.g[v] = *cfg;
    return build_work(b);
}

tsp_status tsp_batch_rk4_steps(tsp_batch *b, int64_t nsteps) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    tsp_status st = build_work(b);
    for (int64_t n = 0; n < nsteps && st == TSP_OK; ++n) {
        st = enqueue_batch_step(b);
        if (st == TSP_OK) batch_advance_clock(b);
    }
    return st;
}

tsp_status tsp_batch_euler_steps(tsp_batch *b, int64_t nsteps) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    tsp_status st = build_work(b);
    for (int64_t n = 0; n < nsteps && st == TSP_OK; ++n) {
        st = enqueue_batch_euler_step(b);
        if (st == TSP_OK) batch_advance_clock(b);
    }
    return st;
}

tsp_status tsp_batch_graph_rk4_steps(tsp_batch *b, int64_t nsteps) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    if (nsteps < 0) return fail(TSP_INVALID_ARG, "nsteps must be >= 0");
    tsp_status st = build_work(b);
    if (st != TSP_OK) return st;
    int slot = -1;
    if (nsteps >= 2 && (st = ensure_batch_graph(b, b->role, &slot)) != TSP_OK) return st;
    int64_t n = 0;
    for (; n + 2 <= nsteps; n += 2) {
        TSP_HIP(hipGraphLaunch(b->graph[slot], b->stream));
        batch_advance_clock(b);
        batch_advance_clock(b);
    }
    if (n < nsteps) return tsp_batch_rk4_steps(b, nsteps - n);
    return TSP_OK;
}

tsp_status tsp_batch_graph_prepare(tsp_batch *b) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    tsp_status st = build_work(b);
    int slot = -1, next[4];
    rk4_next_roles(b->role, next);
    if (st == TSP_OK) st = ensure_batch_graph(b, b->role, &slot);
    if (st == TSP_OK) st = ensure_batch_graph(b, next, &slot);
    return st;
}

tsp_status tsp_batch_time_rk4(tsp_batch *b, int64_t nsteps, int32_t use_graph, double *ms) {
    if (!b || !ms) return fail(TSP_INVALID_ARG, "batch or ms is NULL");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TSP_HIP(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "hipEventCreate");
    }
    tsp_status st = build_work(b);
    if (st == TSP_OK && (e = hipEventRecord(e0, b->stream)) != hipSuccess) st = hip_fail(e, "hipEventRecord");
    if (st == TSP_OK) st = use_graph ? tsp_batch_graph_rk4_steps(b, nsteps) : tsp_batch_rk4_steps(b, nsteps);
    if (st == TSP_OK && (e = hipEventRecord(e1, b->stream)) != hipSuccess) st = hip_fail(e, "hipEventRecord");
    if (st == TSP_OK && (e = hipEventSynchronize(e1)) != hipSuccess) st = hip_fail(e, "hipEventSynchronize");
    float f = 0.f;
    if (st == TSP_OK && (e = hipEventElapsedTime(&f, e0, e1)) != hipSuccess) st = hip_fail(e, "hipEventElapsedTime");
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

tsp_status tsp_batch_sync(tsp_batch *b) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    TSP_HIP(hipStreamSynchronize(b->stream));
    return TSP_OK;
}

tsp_status tsp_batch_get_stream(tsp_batch *b, void **stream) {
    if (!b || !stream) return fail(TSP_INVALID_ARG, "NULL argument");
    *stream = (void *)b->stream;
    return TSP_OK;
}

// ---------------------------------------------------------------------------
// autotuner of the batched kernels
// ---------------------------------------------------------------------------

// cache key: device arch, level or batch, instance count, cells, and the
// (stencil width, KO order) groups present
static std::string batch_tune_key(const tsp_batch *b) {
    hipDeviceProp_t prop{};
    int dev = 0;
    std::string arch = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
    for (char &c : arch)
        if (c == ' ' || c == ':') c = '_';
    int64_t cells = 0;
    for (const tsp_desc &d : b->d) cells += (int64_t)d.n[0] * d.n[1] * d.n[2];
    std::string groups;
    for (int v = 0; v < tsp::kBatchGroups; ++v)
        if (b->group_n[v] > 0)
            groups += "s" + std::to_string(v / tsp::kBatchVariants) + "n" + std::to_string(tsp::kBatchN[v % tsp::kBatchVariants]) +
                      "r" + std::to_string(tsp::kBatchR[v % tsp::kBatchVariants]);
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s|%s|%d|%lld|%s", arch.c_str(), b->level ? "level" : "batch", b->count,
                  (long long)cells, groups.c_str());
    return buf;
}

static bool batch_cache_lookup(const std::string &key, tsp_batch_cfg *cfg) {
    const char *path = std::getenv("TSP_TUNE_CACHE");
    if (!path || !*path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        char k[256];
        tsp_batch_cfg c{};
        double ms = 0;
        if (std::sscanf(line, "%255s %d %d %d %d %d %d %lf", k, &c.tile_y, &c.zchunk, &c.grid_blocks, &c.zalt,
                        &c.halo_ring, &c.narrow_tiles, &ms) == 8 &&
            key == k && validate_batch_cfg(&c) == TSP_OK) {
            *cfg = c;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

static void batch_cache_store(const std::string &key, const tsp_batch_cfg &c, double ms) {
    const char *path = std::getenv("TSP_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d %d %d %d %.6f\n", key.c_str(), c.tile_y, c.zchunk, c.grid_blocks, c.zalt, c.halo_ring,
                 c.narrow_tiles, ms);
    std::fclose(f);
}

// Tunes every work-list group on its own (steps launch that group only): its
// stencil width and KO order set what its best tile shape, z chunk and LDS
// blocking are. The largest group's choice is what tsp_batch_get_kernel
// reports and *chosen receives.
tsp_status tsp_batch_autotune(tsp_batch *b, int32_t steps, tsp_batch_cfg *chosen) {
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    if (steps < 2) steps = 2;
    steps += steps & 1; // whole graph replays
    tsp_status st = build_work(b);
    if (st != TSP_OK) return st;
    const std::string key = batch_tune_key(b);
    // groups present, most cells first
    std::vector<std::pair<int64_t, int>> present;
    for (int v = 0; v < tsp::kBatchGroups; ++v) {
        if (b->group_n[v] == 0) continue;
        int64_t cells = 0;
        for (int i = 0; i < b->count; ++i)
            if (b->h_inst[i].variant == v) cells += (int64_t)b->d[i].n[0] * b->d[i].n[1] * b->d[i].n[2];
        present.push_back({-cells, v});
    }
    std::sort(present.begin(), present.end());
    auto group_key = [&](int v) { return key + "|g" + std::to_string(v); };
    GroupCfgs best = b->gcfg;
    bool cached = true;
    for (auto &pv : present) cached = batch_cache_lookup(group_key(pv.second), &best.g[pv.second]) && cached;
    if (cached) {
        b->gcfg = best;
        b->cfg = best.g[present[0].second];
        if (chosen) *chosen = b->cfg;
        return build_work(b);
    }
    // save the state: roles, clock and every instance's state_old
    int saved_role[4];
    std::memcpy(saved_role, b->role, sizeof(saved_role));
    const int64_t saved_steps = b->steps;
    const std::vector<double> saved_time = b->time;
    const GroupCfgs saved_gcfg = b->gcfg;
    const tsp_batch_cfg saved_cfg = b->cfg;
    int64_t backup_elems = 0;
    for (int i = 0; i < b->count; ++i) backup_elems += b->L[i].elems;
    double *backup = nullptr;
    TSP_HIP(hipMalloc((void **)&backup, sizeof(double) * backup_elems));
    auto copy_old = [&](bool save) -> hipError_t {
        int64_t off = 0;
        for (int i = 0; i < b->count; ++i) {
            double *reg = b->pool + b->reg_off[4 * i + saved_role[TSP_ROLE_OLD]];
            hipError_t e = save ? hipMemcpyAsync(backup + off, reg, sizeof(double) * b->L[i].elems,
                                                 hipMemcpyDeviceToDevice, b->stream)
                                : hipMemcpyAsync(reg, backup + off, sizeof(double) * b->L[i].elems,
                                                 hipMemcpyDeviceToDevice, b->stream);
            if (e != hipSuccess) return e;
            off += b->L[i].elems;
        }
        return hipSuccess;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = copy_old(true);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) st = hip_fail(e, "batch autotune setup");

    std::vector<tsp_batch_cfg> cands;
    for (int ty : {8, 16})
        for (int zc : {0, 64, 128})
            for (int gb : {0, 1024})
                for (int hr = 0; hr < 2; ++hr)
                    for (int nt = 0; nt < 2; ++nt) cands.push_back(tsp_batch_cfg{ty, zc, gb, 1, hr, nt});
    // mean ms per step of group v with configuration c over n graph-replayed steps, from the saved state
    auto time_cfg = [&](int v, const tsp_batch_cfg &c, int n, double *ms) -> tsp_status {
        b->gcfg.g[v] = c;
        tsp_status r = build_work(b);
        if (r == TSP_OK) r = tsp_batch_graph_rk4_steps(b, 2); // capture and warm up
        if (r != TSP_OK) return r;
        if ((e = hipEventRecord(e0, b->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((r = tsp_batch_graph_rk4_steps(b, n)) != TSP_OK) return r;
        if ((e = hipEventRecord(e1, b->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((e = hipEventSynchronize(e1)) != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        float f = 0.f;
        if ((e = hipEventElapsedTime(&f, e0, e1)) != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
        *ms = f / n;
        // restart from the saved state so values stay bounded
        std::memcpy(b->role, saved_role, sizeof(saved_role));
        b->steps = saved_steps;
        b->time = saved_time;
        e = copy_old(false);
        return e == hipSuccess ? TSP_OK : hip_fail(e, "batch autotune restore");
    };
    // per group: every candidate over `steps` steps, then the three fastest over three
    // times as many, each also with zalt 2 (stages 1 and 3 reverse the march)
    std::vector<double> best_ms(tsp::kBatchGroups, 1e300);
    for (size_t gi = 0; gi < present.size() && st == TSP_OK; ++gi) {
        const int v = present[gi].second;
        b->only_group = v;
        std::vector<std::pair<double, int>> ranked;
        for (size_t i = 0; i < cands.size() && st == TSP_OK; ++i) {
            double ms = 0.0;
            st = time_cfg(v, cands[i], steps, &ms);
            ranked.push_back({ms, (int)i});
        }
        if (st != TSP_OK) break;
        std::sort(ranked.begin(), ranked.end());
        for (size_t q = 0; q < std::min<size_t>(3, ranked.size()) && st == TSP_OK; ++q)
            for (int za : {1, 2}) {
                tsp_batch_cfg c = cands[ranked[q].second];
                c.zalt = za;
                double ms = 0.0;
                st = time_cfg(v, c, 3 * steps, &ms);
                if (st != TSP_OK) break;
                if (ms < best_ms[v]) {
                    best_ms[v] = ms;
                    best.g[v] = c;
                }
            }
        b->gcfg.g[v] = best.g[v]; // the later groups are timed beside nothing else, but build with it
    }
    b->only_group = -1;
    // restore (the state is already restored after every candidate)
    std::memcpy(b->role, saved_role, sizeof(saved_role));
    b->steps = saved_steps;
    b->time = saved_time;
    e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess && st == TSP_OK) st = hip_fail(e, "batch autotune");
    (void)hipFree(backup);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st == TSP_OK) {
        b->gcfg = best;
        b->cfg = best.g[present[0].second];
    } else {
        b->gcfg = saved_gcfg;
        b->cfg = saved_cfg;
    }
    tsp_status st2 = build_work(b);
    if (st != TSP_OK) return st;
    if (st2 != TSP_OK) return st2;
    for (auto &pv : present) batch_cache_store(group_key(pv.second), best.g[pv.second], best_ms[pv.second]);
    if (chosen) *chosen = b->cfg;
    return TSP_OK;
}

// ---------------------------------------------------------------------------
// a level split into boxes (SingleLevelIntegrator.hpp:47-49,172-176)
// ---------------------------------------------------------------------------

// Cut [0, len) into pieces of at most `bs` cells: the fewest pieces, sizes as
// equal as possible (the larger ones first). Returns the piece starts and len.
static std::vector<int> chop_axis(int len, int bs) {
    const int nb = (len + bs - 1) / bs;
    const int sz = len / nb, extra = len % nb;
    std::vector<int> cuts{0};
    for (int p = 0; p < nb; ++p) cuts.push_back(cuts.back() + sz + (p < extra ? 1 : 0));
    return cuts;
}

tsp_status tsp_level_create(const tsp_desc *d, const int32_t *block, void *stream, tsp_batch **out) {
    if (!out) return fail(TSP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!block) return fail(TSP_INVALID_ARG, "block size is NULL");
    tsp_status st = validate(d);
    if (st != TSP_OK) return st;
    if (d->bc == TSP_BC_CUSTOM) return fail(TSP_UNSUPPORTED, "multi-box levels carry NoOp and Dirichlet boundaries only");
    if (tsp::batch_system_slot(d->system) < 0)
        return fail(TSP_UNSUPPORTED, "multi-box levels run the (phi,Pi), (phi,u) and NonLinear systems");
    for (int a = 0; a < 3; ++a)
        if (block[a] < 1) return fail(TSP_INVALID_ARG, "block size entries must be >= 1");
    std::vector<int> cuts[3];
    for (int a = 0; a < 3; ++a) cuts[a] = chop_axis(d->n[a], block[a]);
    const int nb[3] = {(int)cuts[0].size() - 1, (int)cuts[1].size() - 1, (int)cuts[2].size() - 1};
    const long long nboxes = (long long)nb[0] * nb[1] * nb[2];
    if (nboxes > 65535) return fail(TSP_UNSUPPORTED, "a level holds at most 65535 boxes");
    std::vector<tsp_desc> descs;
    std::vector<int32_t> offs;
    for (int bz = 0; bz < nb[2]; ++bz)
        for (int by = 0; by < nb[1]; ++by)
            for (int bx = 0; bx < nb[0]; ++bx) { // instance b = bx + nb0 * (by + nb1 * bz)
                tsp_desc bd = *d;
                const int lo[3] = {cuts[0][bx], cuts[1][by], cuts[2][bz]};
                const int hi[3] = {cuts[0][bx + 1], cuts[1][by + 1], cuts[2][bz + 1]};
                for (int a = 0; a < 3; ++a) {
                    bd.n[a] = hi[a] - lo[a];
                    bd.periodic[a] = 0; // the exchange supplies neighbours and periodic images
                    offs.push_back(lo[a]);
                }
                descs.push_back(bd);
            }
    tsp_batch *b = nullptr;
    st = create_batch(descs.data(), (int32_t)nboxes, stream, &b, offs.data(), d->n);
    if (st != TSP_OK) return st;
    b->level = true;
    b->ldesc = *d;
    b->box_lo = offs;
    std::vector<int32_t> axis_box;
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < nb[a]; ++q)
            for (int x = cuts[a][q]; x < cuts[a][q + 1]; ++x) axis_box.push_back(q);
    hipError_t e = hipMalloc((void **)&b->d_axis_box, sizeof(int32_t) * axis_box.size());
    if (e == hipSuccess)
        e = hipMemcpy(b->d_axis_box, axis_box.data(), sizeof(int32_t) * axis_box.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        tsp_batch_destroy(b);
        return hip_fail(e, "level box table");
    }
    tsp::LevelArgs &L = b->largs;
    L.inst = b->d_inst;
    L.axis_box = b->d_axis_box;
    for (int a = 0; a < 3; ++a) {
        L.dom[a] = d->n[a];
        L.nbox[a] = nb[a];
        L.periodic[a] = d->periodic[a];
    }
    L.dirichlet = d->bc == TSP_BC_DIRICHLET;
    L.G = d->ghost;
    L.nc = system_ncomp(d->system);
    L.nboxes = (int32_t)nboxes;
    for (int c = 0; c < TSP_MAX_COMP; ++c) L.target[c] = d->target[c];
    for (int i = 0; i < b->count; ++i) {
        const tsp_desc &bd = b->d[i];
        const long long g = (long long)(bd.n[0] + 2 * L.G) * (bd.n[1] + 2 * L.G) * (bd.n[2] + 2 * L.G) -
                            (long long)bd.n[0] * bd.n[1] * bd.n[2];
        b->max_shell = std::max(b->max_shell, g);
    }
    *out = b;
    return TSP_OK;
}

tsp_status tsp_level_boxes(tsp_batch *b, int32_t *lo_hi) {
    if (!b || !lo_hi) return fail(TSP_INVALID_ARG, "level or output is NULL");
    if (!b->level) return fail(TSP_INVALID_ARG, "not a level (tsp_level_create)");
    for (int i = 0; i < b->count; ++i)
        for (int a = 0; a < 3; ++a) {
            lo_hi[6 * i + a] = b->box_lo[3 * i + a];
            lo_hi[6 * i + 3 + a] = b->box_lo[3 * i + a] + b->d[i].n[a] - 1;
        }
    return TSP_OK;
}

static tsp_status level_copy(tsp_batch *b, double *buf, int32_t on_device, bool scatter) {
    if (!b || !buf) return fail(TSP_INVALID_ARG, "level or buffer is NULL");
    if (!b->level) return fail(TSP_INVALID_ARG, "not a level (tsp_level_create)");
    const size_t bytes =
        sizeof(double) * (size_t)b->ldesc.n[0] * b->ldesc.n[1] * b->ldesc.n[2] * system_ncomp(b->ldesc.system);
    double *dbuf = buf;
    double *staging = nullptr;
    if (!on_device) {
        TSP_HIP(hipMalloc((void **)&staging, bytes));
        dbuf = staging;
    }
    hipError_t e = hipSuccess;
    if (scatter && staging) e = hipMemcpyAsync(staging, buf, bytes, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = tsp::launch_level_copy(b->largs, b->role[TSP_ROLE_OLD], dbuf, scatter, b->stream);
    if (e == hipSuccess && !scatter && staging) e = hipMemcpyAsync(buf, staging, bytes, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (staging) (void)hipFree(staging);
    if (e != hipSuccess) return hip_fail(e, scatter ? "level set_state" : "level get_state");
    return TSP_OK;
}

tsp_status tsp_level_set_state(tsp_batch *b, const double *src, int32_t on_device) {
    if (!b || !b->level) return fail(TSP_INVALID_ARG, "not a level (tsp_level_create)");
    // initialise(): the clock and the register roles restart
    std::memcpy(b->role, kIdentityRoles, sizeof(b->role));
    b->steps = 0;
    b->time.assign(b->count, 0.0);
    return level_copy(b, const_cast<double *>(src), on_device, true);
}

tsp_status tsp_level_get_state(tsp_batch *b, double *dst, int32_t on_device) {
    return level_copy(b, dst, on_device, false);
}

tsp_status tsp_level_init_expr(tsp_batch *b, const tsp_expr_op *ops, const int32_t *comp_start,
                               const int32_t index_lo[3]) {
    if (!b || !b->level) return fail(TSP_INVALID_ARG, "not a level (tsp_level_create)");
    const int nc = desc_ncomp(b->ldesc);
    if (tsp_status st = check_expr_programs(ops, comp_start, nc); st != TSP_OK) return st;
    std::memcpy(b->role, kIdentityRoles, sizeof(b->role));
    b->steps = 0;
    b->time.assign(b->count, 0.0);
    std::vector<tsp::InitBox> boxes(b->count);
    for (int i = 0; i < b->count; ++i) {
        const Layout &L = b->L[i];
        boxes[i] = tsp::InitBox{batch_origin(b, i, b->role[TSP_ROLE_OLD]), L.sj, L.sk, L.sc,
                                {b->d[i].n[0], b->d[i].n[1], b->d[i].n[2]},
                                {b->box_lo[3 * i], b->box_lo[3 * i + 1], b->box_lo[3 * i + 2]}};
    }
    return run_init_expr(boxes, b->ldesc, ops, comp_start, nc, index_lo, b->stream);
}

} // extern "C"

// ===========================================================================
// Discrete adjoint of a batch: every instance's receiver loss gradient at once
// ===========================================================================
//
// The same recursion as the single-box adjoint (capi.hip), with every launch
// covering all instances: forward steps are the batch's RK4 steps followed by
// one gather of every instance's receivers, reverse steps one inject and four
// reverse-stage launches per variant group on the adjoint registers. Each
// instance's trace and gradient are bit-identical to its single-box adjoint.

struct tsp_batch_adjoint {
    tsp_batch *b = nullptr;
    int nrec = 0, nsteps = 0, K = 0, nck = 0;
    int64_t step0 = 0;
    std::vector<double> time0;
    std::vector<int64_t> state_off; // element offset of instance i in a dense [instance][state] array
    int64_t state_elems = 0;        // all instances
    double *apool = nullptr;        // adjoint registers, the forward pool's layout
    std::vector<tsp::BatchInstance> h_ainst;
    tsp::BatchInstance *d_ainst = nullptr;
    int *d_rec = nullptr;           // [instance][receiver] offsets of component 0 from the register origin
    double *d_trace = nullptr;      // [instance][step][receiver]
    double *d_dtrace = nullptr;
    int *d_counter = nullptr;       // one trace row counter per instance
    double *d_ckpt = nullptr;       // [slot][instance state]
    int arole[4] = {0, 1, 2, 3};
    int roles0[4] = {0, 1, 2, 3};  // the batch's register roles when the forward pass started
    int groles0[4] = {-1, -1, -1, -1}; // roles0 the forward graphs were captured for
    hipGraphExec_t fgraph[2] = {nullptr, nullptr}; // two (step, gather) pairs per forward parity
    hipGraphExec_t bgraph[2] = {nullptr, nullptr}; // two (inject, reverse step) pairs per backward parity
    uint64_t gwork = ~0ull; // work-list generation the graphs and the adjoint instance table were made for
};

namespace {

void drop_batch_adjoint_graphs(tsp_batch_adjoint *a) {
    for (auto *g : {&a->fgraph[0], &a->fgraph[1], &a->bgraph[0], &a->bgraph[1]})
        if (*g) {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

// Graphs hold the batch's work list, and the adjoint instance table the z
// chunk of every instance; both follow a rebuilt work list (every change of a
// group's configuration rebuilds it).
tsp_status batch_adjoint_check(tsp_batch_adjoint *a) {
    tsp_status st = build_work(a->b);
    if (st != TSP_OK) return st;
    if (a->gwork != a->b->work_gen) {
        drop_batch_adjoint_graphs(a);
        for (int i = 0; i < a->b->count; ++i) a->h_ainst[i].zchunk = a->b->h_inst[i].zchunk;
        TSP_HIP(hipMemcpyAsync(a->d_ainst, a->h_ainst.data(), sizeof(tsp::BatchInstance) * a->b->count,
                               hipMemcpyHostToDevice, a->b->stream));
        TSP_HIP(hipStreamSynchronize(a->b->stream));
        a->gwork = a->b->work_gen;
    }
    return TSP_OK;
}

tsp_status enqueue_batch_fwd_gather(tsp_batch_adjoint *a) {
    tsp_batch *b = a->b;
    tsp_status st = enqueue_batch_step(b);
    if (st != TSP_OK) return st;
    TSP_HIP(tsp::launch_batch_gather_trace(b->d_inst, b->count, b->role[TSP_ROLE_OLD], a->d_rec, a->nrec, a->d_trace,
                                           a->nsteps, a->d_counter, b->stream));
    return TSP_OK;
}

// lambda <- G^T lambda for every instance (lambda in adjoint role OLD)
tsp_status enqueue_batch_adjoint_step(tsp_batch_adjoint *a) {
    tsp_batch *b = a->b;
    int &OLD = a->arole[TSP_ROLE_OLD], &NEW = a->arole[TSP_ROLE_NEW], &ARG = a->arole[TSP_ROLE_ARG],
        &ARGN = a->arole[TSP_ROLE_ARG_NEXT];
    for (int j = 3; j >= 0; --j) {
        if (j < 3) std::swap(ARG, ARGN);
        for (int v = 0; v < tsp::kBatchVariants; ++v) {
            if (b->group_n[v] == 0) continue;
            tsp::BatchArgs g{};
            g.inst = a->d_ainst;
            g.work = b->d_work + b->group_off[v];
            g.nwork = b->group_n[v];
            g.stage = j;
            g.src = j == 3 ? OLD : ARG;
            g.init = OLD;
            g.acc = NEW;
            g.argn = j > 0 ? ARGN : -1;
            const tsp_batch_cfg &gc = b->gcfg.g[v];
            g.zalt = gc.zalt;
            g.ghost_mode = 0;
            const int grid = gc.grid_blocks > 0 ? std::min(gc.grid_blocks, g.nwork) : g.nwork;
            TSP_HIP(tsp::launch_batched_phipi_adjoint(v, g, gc.tile_y, grid, b->max_abl_T, gc.halo_ring,
                                                      b->stream));
        }
    }
    std::swap(OLD, NEW);
    return TSP_OK;
}

tsp_status enqueue_batch_inject_adjoint(tsp_batch_adjoint *a) {
    TSP_HIP(tsp::launch_batch_inject_trace(a->d_ainst, a->b->count, a->arole[TSP_ROLE_OLD], a->d_rec, a->nrec,
                                           a->d_dtrace, a->nsteps, a->d_counter, a->b->stream));
    return enqueue_batch_adjoint_step(a);
}

template <class F> tsp_status batch_capture_pair(hipStream_t stream, hipGraphExec_t *slot, int *roles, F &&body) {
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

tsp_status batch_run_forward_range(tsp_batch_adjoint *a, int n0, int n1) {
    tsp_batch *b = a->b;
    int n = n0;
    while (n < n1) {
        if (n1 - n >= 2) {
            const int parity = (int)((b->steps - a->step0) & 1); // roles are roles0 after an even count
            tsp_status st =
                batch_capture_pair(b->stream, &a->fgraph[parity], b->role, [&] { return enqueue_batch_fwd_gather(a); });
            if (st != TSP_OK) return st;
            TSP_HIP(hipGraphLaunch(a->fgraph[parity], b->stream));
            batch_advance_clock(b);
            batch_advance_clock(b);
            n += 2;
        } else {
            tsp_status st = enqueue_batch_fwd_gather(a);
            if (st != TSP_OK) return st;
            batch_advance_clock(b);
            n += 1;
        }
    }
    return TSP_OK;
}

tsp_status batch_run_backward_range(tsp_batch_adjoint *a, int n1, int n0, int *bparity) {
    int n = n1;
    while (n > n0) {
        if (n - n0 >= 2) {
            tsp_status st = batch_capture_pair(a->b->stream, &a->bgraph[*bparity], a->arole,
                                               [&] { return enqueue_batch_inject_adjoint(a); });
            if (st != TSP_OK) return st;
            TSP_HIP(hipGraphLaunch(a->bgraph[*bparity], a->b->stream));
            n -= 2;
        } else {
            tsp_status st = enqueue_batch_inject_adjoint(a);
            if (st != TSP_OK) return st;
            *bparity ^= 1;
            n -= 1;
        }
    }
    return TSP_OK;
}

int batch_ckpt_slot_of_step(const tsp_batch_adjoint *a, int step) {
    if (step == a->nsteps) return a->nck - 1;
    return step % a->K == 0 ? step / a->K : -1;
}

tsp_status batch_save_checkpoint(tsp_batch_adjoint *a, int slot) {
    tsp_batch *b = a->b;
    for (int i = 0; i < b->count; ++i) {
        const tsp_desc &d = b->d[i];
        const Layout &L = b->L[i];
        TSP_HIP(tsp::launch_pack(a->d_ckpt + slot * a->state_elems + a->state_off[i],
                                 batch_origin(b, i, b->role[TSP_ROLE_OLD]), d.n[0], d.n[1], d.n[2],
                                 system_ncomp(d.system), (int)L.sj, (int)L.sk, (int)L.sc, b->stream));
    }
    return TSP_OK;
}

// checkpoint `slot` (the state `step` steps into the forward pass) back into the batch
tsp_status batch_restore_checkpoint(tsp_batch_adjoint *a, int slot, int step, const std::vector<double> &time) {
    tsp_batch *b = a->b;
    b->steps = a->step0 + step;
    if (step & 1) rk4_next_roles(a->roles0, b->role);
    else std::memcpy(b->role, a->roles0, sizeof(b->role));
    b->time = time;
    for (int i = 0; i < b->count; ++i) {
        const tsp_desc &d = b->d[i];
        const Layout &L = b->L[i];
        TSP_HIP(tsp::launch_unpack(batch_origin(b, i, b->role[TSP_ROLE_OLD]),
                                   a->d_ckpt + slot * a->state_elems + a->state_off[i], d.n[0], d.n[1], d.n[2],
                                   system_ncomp(d.system), (int)L.sj, (int)L.sk, (int)L.sc, b->stream));
    }
    return TSP_OK;
}

} // namespace

extern "C" {

tsp_status tsp_batch_adjoint_create(tsp_batch *b, const int32_t *receivers, int32_t nrec, int32_t nsteps,
                                    int32_t checkpoint_every, tsp_batch_adjoint **out) {
    if (!out) return fail(TSP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!b) return fail(TSP_INVALID_ARG, "batch is NULL");
    if (nrec < 1 || !receivers) return fail(TSP_INVALID_ARG, "need at least one receiver");
    if (nsteps < 1) return fail(TSP_INVALID_ARG, "nsteps must be >= 1");
    if (checkpoint_every < 1) return fail(TSP_INVALID_ARG, "checkpoint_every must be >= 1");
    if (b->level) return fail(TSP_UNSUPPORTED, "the batched adjoint runs batches of single boxes, not a level");
    for (int i = 0; i < b->count; ++i)
        if (b->d[i].system != TSP_SYS_WAVE_PHIPI)
            return fail(TSP_UNSUPPORTED, "instance " + std::to_string(i) +
                                             ": the discrete adjoint is available for the (phi,Pi) system");
    std::vector<int> off((size_t)b->count * nrec);
    for (int i = 0; i < b->count; ++i)
        for (int r = 0; r < nrec; ++r) {
            const int x = receivers[3 * r], y = receivers[3 * r + 1], z = receivers[3 * r + 2];
            const tsp_desc &d = b->d[i];
            if (x < 0 || x >= d.n[0] || y < 0 || y >= d.n[1] || z < 0 || z >= d.n[2])
                return fail(TSP_INVALID_ARG, "receiver " + std::to_string(r) + " lies outside the valid box of instance " +
                                                 std::to_string(i));
            off[(size_t)i * nrec + r] = (int)(x + y * b->L[i].sj + z * b->L[i].sk); // component 0 (phi)
        }
    tsp_batch_adjoint *a = new (std::nothrow) tsp_batch_adjoint();
    if (!a) return fail(TSP_OOM, "host allocation failed");
    auto cleanup = [&](tsp_status e) {
        tsp_batch_adjoint_destroy(a);
        return e;
    };
    a->b = b;
    a->nrec = nrec;
    a->nsteps = nsteps;
    a->K = checkpoint_every;
    a->nck = nsteps / checkpoint_every + 1 + (nsteps % checkpoint_every != 0 ? 1 : 0);
    for (int i = 0; i < b->count; ++i) {
        a->state_off.push_back(a->state_elems);
        a->state_elems += (int64_t)b->d[i].n[0] * b->d[i].n[1] * b->d[i].n[2] * system_ncomp(b->d[i].system);
    }
    const size_t trace_elems = (size_t)b->count * nsteps * nrec;
    hipError_t e = hipMalloc((void **)&a->d_rec, sizeof(int) * off.size());
    if (e == hipSuccess) e = hipMemcpy(a->d_rec, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_trace, sizeof(double) * trace_elems);
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_dtrace, sizeof(double) * trace_elems);
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_counter, sizeof(int) * b->count);
    if (e == hipSuccess) e = hipMalloc((void **)&a->d_ckpt, sizeof(double) * a->state_elems * a->nck);
    // adjoint registers: the forward pool's layout, zero everywhere (ghosts stay zero:
    // homogeneous boundary values)
    if (e == hipSuccess) e = hipMalloc((void **)&a->apool, sizeof(double) * b->pool_elems);
    if (e == hipSuccess) e = hipMemsetAsync(a->apool, 0, sizeof(double) * b->pool_elems, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(a->d_trace, 0, sizeof(double) * trace_elems, b->stream);
    if (e == hipSuccess) {
        a->h_ainst = b->h_inst;
        for (int i = 0; i < b->count; ++i) {
            tsp::BatchInstance &I = a->h_ainst[i];
            for (int r = 0; r < 4; ++r) I.reg[r] = I.reg[r] - b->pool + a->apool;
            for (int c = 0; c < TSP_MAX_COMP; ++c) I.target[c] = 0.0;
        }
        e = hipMalloc((void **)&a->d_ainst, sizeof(tsp::BatchInstance) * b->count);
    }
    if (e == hipSuccess)
        e = hipMemcpy(a->d_ainst, a->h_ainst.data(), sizeof(tsp::BatchInstance) * b->count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e == hipErrorOutOfMemory) return cleanup(fail(TSP_OOM, "adjoint buffers do not fit in device memory"));
    if (e != hipSuccess) return cleanup(hip_fail(e, "batch adjoint allocation"));
    *out = a;
    return TSP_OK;
}

tsp_status tsp_batch_adjoint_destroy(tsp_batch_adjoint *a) {
    if (!a) return TSP_OK;
    if (a->b && a->b->stream) (void)hipStreamSynchronize(a->b->stream);
    drop_batch_adjoint_graphs(a);
    for (void *p : {(void *)a->d_rec, (void *)a->d_trace, (void *)a->d_dtrace, (void *)a->d_counter, (void *)a->d_ckpt,
                    (void *)a->apool, (void *)a->d_ainst})
        if (p) (void)hipFree(p);
    delete a;
    return TSP_OK;
}

tsp_status tsp_batch_adjoint_forward(tsp_batch_adjoint *a) {
    if (!a) return fail(TSP_INVALID_ARG, "adjoint is NULL");
    tsp_batch *b = a->b;
    tsp_status st = batch_adjoint_check(a);
    if (st != TSP_OK) return st;
    a->step0 = b->steps;
    a->time0 = b->time;
    std::memcpy(a->roles0, b->role, sizeof(a->roles0));
    if (std::memcmp(a->groles0, a->roles0, sizeof(a->roles0)) != 0) { // forward graphs belong to other roles
        drop_batch_adjoint_graphs(a);
        std::memcpy(a->groles0, a->roles0, sizeof(a->roles0));
    }
    if ((st = batch_save_checkpoint(a, 0)) != TSP_OK) return st;
    TSP_HIP(tsp::launch_batch_set_counter(a->d_counter, b->count, 0, b->stream));
    int n = 0;
    while (n < a->nsteps && st == TSP_OK) {
        const int e = std::min(a->nsteps, (n / a->K + 1) * a->K);
        st = batch_run_forward_range(a, n, e);
        n = e;
        if (st == TSP_OK) {
            const int slot = batch_ckpt_slot_of_step(a, n);
            if (slot >= 0) st = batch_save_checkpoint(a, slot);
        }
    }
    return st;
}

tsp_status tsp_batch_adjoint_trace(tsp_batch_adjoint *a, double *dst, int32_t on_device) {
    if (!a || !dst) return fail(TSP_INVALID_ARG, "adjoint or destination is NULL");
    TSP_HIP(hipStreamSynchronize(a->b->stream));
    TSP_HIP(hipMemcpy(dst, a->d_trace, sizeof(double) * (size_t)a->b->count * a->nsteps * a->nrec,
                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
    return TSP_OK;
}

tsp_status tsp_batch_adjoint_backward(tsp_batch_adjoint *a, const double *dtrace, int32_t dtrace_on_device,
                                      double *grad, int32_t grad_on_device, int32_t recompute) {
    if (!a || !dtrace || !grad) return fail(TSP_INVALID_ARG, "adjoint, dtrace or grad is NULL");
    tsp_batch *b = a->b;
    tsp_status st = batch_adjoint_check(a);
    if (st != TSP_OK) return st;
    TSP_HIP(hipMemcpyAsync(a->d_dtrace, dtrace, sizeof(double) * (size_t)b->count * a->nsteps * a->nrec,
                           dtrace_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, b->stream));
    std::memcpy(a->arole, kIdentityRoles, sizeof(a->arole));
    for (int i = 0; i < b->count; ++i) // lambda = 0 in adjoint register 0 (ghosts stay zero)
        TSP_HIP(tsp::launch_zero(a->apool + b->reg_off[4 * i], b->L[i].elems, b->stream));
    int bparity = 0;
    if (!recompute) {
        TSP_HIP(tsp::launch_batch_set_counter(a->d_counter, b->count, a->nsteps - 1, b->stream));
        st = batch_run_backward_range(a, a->nsteps, 0, &bparity);
    } else {
        // segments from the last to the first: restore the segment's checkpoint,
        // recompute its forward states and trace, then reverse it
        const std::vector<double> saved_time = b->time;
        const int nseg = (a->nsteps + a->K - 1) / a->K;
        std::vector<std::vector<double>> seg_time(nseg + 1, a->time0);
        for (int m = 1; m <= nseg; ++m) { // every instance's clock at each segment start, accumulated as the steps do
            seg_time[m] = seg_time[m - 1];
            const int len = std::min(a->K, a->nsteps - (m - 1) * a->K);
            for (int i = 0; i < b->count; ++i)
                for (int q = 0; q < len; ++q) seg_time[m][i] += b->d[i].dt;
        }
        for (int m = nseg - 1; m >= 0 && st == TSP_OK; --m) {
            const int n0 = m * a->K, n1 = std::min(a->nsteps, n0 + a->K);
            st = batch_restore_checkpoint(a, m, n0, seg_time[m]);
            if (st == TSP_OK) {
                hipError_t e = tsp::launch_batch_set_counter(a->d_counter, b->count, n0, b->stream);
                if (e != hipSuccess) st = hip_fail(e, "set counter");
            }
            if (st == TSP_OK) st = batch_run_forward_range(a, n0, n1);
            if (st == TSP_OK) {
                hipError_t e = tsp::launch_batch_set_counter(a->d_counter, b->count, n1 - 1, b->stream);
                if (e != hipSuccess) st = hip_fail(e, "set counter");
            }
            if (st == TSP_OK) st = batch_run_backward_range(a, n1, n0, &bparity);
        }
        // leave the batch at the end of the forward pass
        if (st == TSP_OK) st = batch_restore_checkpoint(a, a->nck - 1, a->nsteps, saved_time);
    }
    if (st != TSP_OK) return st;
    double *dst = grad;
    double *staging = nullptr;
    if (!grad_on_device) {
        TSP_HIP(hipMalloc((void **)&staging, sizeof(double) * a->state_elems));
        dst = staging;
    }
    hipError_t e = hipSuccess;
    for (int i = 0; i < b->count && e == hipSuccess; ++i) {
        const tsp_desc &d = b->d[i];
        const Layout &L = b->L[i];
        e = tsp::launch_pack(dst + a->state_off[i], a->h_ainst[i].reg[a->arole[TSP_ROLE_OLD]], d.n[0], d.n[1], d.n[2],
                             system_ncomp(d.system), (int)L.sj, (int)L.sk, (int)L.sc, b->stream);
    }
    if (e == hipSuccess && staging)
        e = hipMemcpyAsync(grad, staging, sizeof(double) * a->state_elems, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (staging) (void)hipFree(staging);
    if (e != hipSuccess) return hip_fail(e, "batch adjoint gradient");
    return TSP_OK;
}

} // extern "C"
