/*
 * treesap_hip.h -- C ABI of the MI355X (gfx950) explicit time-stepping library.
 *
 * The reference (TreESaP) has no C/FFI boundary: its plug-in API is the
 * compile-time C++ concept set of pde-model/PDEModelConcepts.hpp:44-139 and
 * the integrator lifecycle of Integrator.hpp:75-106 / SingleLevelIntegrator.hpp.
 * Each entry point below replaces one piece of that lifecycle; the comment on
 * each names the reference interface it stands in for. The C++ drop-in headers
 * under include/treesap/ (PDEModel, SingleLevelRK4Integrator, ...) call these
 * entry points, as do the Python binding and the tests.
 *
 * Conventions
 *   - Every function returns a tsp_status; on failure tsp_last_error() returns
 *     a thread-local message. Nothing throws across this boundary.
 *   - A solver owns its device memory and is bound to one HIP stream. Calls on
 *     one solver must not overlap (no internal locking). Host buffers passed in
 *     are borrowed for the duration of the call only.
 *   - State arrays exchanged through tsp_set_state/tsp_get_state hold the valid
 *     region only, FP64, component-slowest then z, y, x fastest:
 *       a[((c * nz + k) * ny + j) * nx + i].
 *   - All arithmetic is IEEE FP64 without FMA contraction, in the reference's
 *     operation order (amrex::Real is double).
 */
#ifndef TREESAP_HIP_H
#define TREESAP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSP_MAX_COMP 12
#define TSP_VERSION "0.1.0"

typedef enum tsp_status {
    TSP_OK = 0,
    TSP_INVALID_ARG = 1,
    TSP_OOM = 2,
    TSP_HIP_ERROR = 3,
    TSP_UNSUPPORTED = 4
} tsp_status;

/* Registered right-hand sides (the RHS functor is a compile-time type in the
 * reference, RHSConcept at pde-model/PDEModelConcepts.hpp:44-64; the library
 * carries device kernels for these systems). */
typedef enum tsp_system {
    TSP_SYS_WAVE_PHIPI = 0, /* (phi, Pi): phi_t = Pi, Pi_t = sum_a D_aa phi           */
    TSP_SYS_WAVE_PHIU = 1,  /* (phi, u_a): examples/01-wave-eqn/WaveEqnRHS.hpp:16-50    */
    TSP_SYS_NONLINEAR = 2,  /* (phi, pi): examples/02-temporaries/NonLinearRHS.hpp:16-71 */
    TSP_SYS_TENSOR_ADM = 3, /* (h_ij, K_ij) linearised ADM, 12 comps, SymRank2 aliases   */
    TSP_SYS_CUSTOM = 4      /* any RHSConcept functor: the caller's stage kernel evaluates it
                               (tsp_rk4_steps_custom); tsp_desc.ncomp components; the KO and
                               ABL decorators belong to the caller's functor (ko_order and
                               abl_thickness must be 0) */
} tsp_system;

/* Physical boundary functor on non-periodic faces (BCConcept,
 * PDEModelConcepts.hpp:70-90): NoOpBC (bcs/NoOpBC.hpp:22-47) leaves ghost
 * memory untouched; DirichletBC (bcs/DirichletBC.hpp:21-66) writes target[];
 * TSP_BC_CUSTOM: any other functor, applied by the caller's callback
 * (tsp_set_bc_callback) after the library's periodic fill. */
typedef enum tsp_bc { TSP_BC_NOOP = 0, TSP_BC_DIRICHLET = 1, TSP_BC_CUSTOM = 2 } tsp_bc;

/* ABLProfileConcept (rhs-decorators/ABLDecorator.hpp:24-29) instances. */
typedef enum tsp_profile { TSP_PROFILE_LINEAR = 0, TSP_PROFILE_QUADRATIC = 1, TSP_PROFILE_CUBIC = 2 } tsp_profile;

/* Register roles (SingleLevelIntegrator.hpp:182-183 and
 * integrators/SingleLevelRK4Integrator.hpp:162-166). */
typedef enum tsp_role { TSP_ROLE_OLD = 0, TSP_ROLE_NEW = 1, TSP_ROLE_ARG = 2, TSP_ROLE_ARG_NEXT = 3 } tsp_role;

/* One PDE model instance: the runtime content of PDEModel
 * (pde-model/PDEModel.hpp:70-77), KODecorator (KODecorator.hpp:32-74),
 * ABLPDEModel/ABLDecorator::ABLData (ABLPDEModel.hpp:47-50,
 * ABLDecorator.hpp:58-77) and amrex::Geometry. */
typedef struct tsp_desc {
    int32_t system;         /* tsp_system */
    int32_t n[3];           /* valid cells per axis (Geometry::Domain) */
    int32_t stencil_n;      /* derivative half-width (treesap::stencil ngrow template) */
    int32_t ghost;          /* ghost width G = RHS::ngrow, >= stencil_n and >= ko_order */
    int32_t ko_order;       /* KODecorator order r; 0 = no KO decorator */
    int32_t bc;             /* tsp_bc */
    double ko_strength;     /* KODecorator::strength */
    double prob_lo[3];      /* RealBox lo */
    double prob_hi[3];      /* RealBox hi */
    int32_t periodic[3];    /* Geometry is_periodic */
    int32_t abl_thickness;  /* ABLData::thickness in cells; 0 = no ABL decorator */
    int32_t abl_profile;    /* tsp_profile */
    int32_t ncomp;          /* component count of a TSP_SYS_CUSTOM model (ignored otherwise) */
    double abl_max;         /* ABLData::max_strength */
    double target[TSP_MAX_COMP]; /* ABLData::target_values == DirichletBC::boundary_values */
    double gamma;           /* NonLinearRHS::gamma (system 2 only) */
    double dt;              /* PDEModel::dt_hint() */
} tsp_desc;

/* Stage-kernel configuration chosen by tsp_autotune or set by hand. Every
 * configuration produces bit-identical results; only the speed differs. */
typedef struct tsp_kernel_cfg {
    int32_t ghost_mode; /* 0: virtual ghosts (no fill pass); 1: FillPatch pass then plain reads */
    int32_t tile_y;     /* rows per workgroup: 2, 4, 8 or 16 with 64-wide tiles (one wave per row),
                           or 32 for 32 x 32 tiles (two rows per wave). The tensor system's block
                           shape: 8 -> 16 x 8 x 2 cells, 4 -> 8 x 8 x 4, 16 -> 16 x 8 x 4 (512
                           threads), others 16 x 8 x 2 */
    int32_t zchunk;     /* z planes marched per workgroup (>= 1); the tensor system: z tiles per
                           slab of its XCD-wise block order */
    int32_t xcd_remap;  /* 1: map consecutive tiles onto one XCD (shared L2) */
    int32_t zalt;       /* 1: odd z chunks march downwards (chunk-boundary planes read together);
                           2: as 1, and RK4 stages 1 and 3 reverse every direction (a stage
                           starts at the planes the previous stage wrote last) */
    int32_t halo_ring;  /* 1: tile_y 8/16 load each plane's x/y halo with its z-queue head and keep
                           it in LDS until use (falls back to 0 where the LDS does not fit) */
    int32_t stage_pairs; /* 1: RK4 stages (0,1) and (2,3) each run as one launch on 32 x 32 tiles
                            (the first stage of a pair also on an H-wide ring of neighbour cells;
                            tile_y, zalt and halo_ring do not apply). (phi, Pi) system with KO
                            dissipation, ghost_mode 0, registers under 4 GiB and periodic axes of
                            at least 2*max(stencil_n, ko_order) cells; otherwise tsp_set_kernel
                            returns TSP_UNSUPPORTED. The intermediate stage registers (roles arg
                            and arg_next) are not written in this mode. */
} tsp_kernel_cfg;

typedef struct tsp_solver tsp_solver;

/* Thread-local message describing the last failure on this thread. */
const char *tsp_last_error(void);
const char *tsp_version(void);

/* Fill a descriptor with defaults: periodic unit cube, no decorators. */
tsp_status tsp_desc_default(tsp_desc *d, int32_t system, int32_t n);

/* Replaces SingleLevelRK4Integrator(model, block_size)
 * (integrators/SingleLevelRK4Integrator.hpp:65-67, SingleLevelIntegrator.hpp:47-49):
 * validates the model, allocates the four registers (state_old, state_new,
 * arg, arg_next) in device memory (the tensor system with KO dissipation also
 * a six-component buffer for its K components' KO sums, written and read
 * within each stage). `stream` is a hipStream_t or NULL (the
 * library then creates its own default-priority stream). The descriptor is
 * validated before any HIP call; with an absorbing layer the device's
 * per-workgroup LDS is then queried, since the stage kernels keep the layer's
 * rate table (3 T^2 + 1 doubles) in LDS (TSP_UNSUPPORTED if it does not fit).
 * Stage kernels address registers under 4 GiB through buffer descriptors and
 * larger ones (up to the 2^31-element limit) through 64-bit pointers; setting
 * the environment variable TSP_POINTER_ACCESS before tsp_create selects the
 * pointer-access kernel for any size (results are bit-identical). */
tsp_status tsp_create(const tsp_desc *d, void *stream, tsp_solver **out);
tsp_status tsp_destroy(tsp_solver *s);

/* Replaces SingleLevelIntegrator::initialise_impl (SingleLevelIntegrator.hpp:75-93):
 * the initial condition is evaluated by the caller and uploaded into state_old.
 * on_device != 0 means `src` is a device pointer. Resets time and step count. */
tsp_status tsp_set_state(tsp_solver *s, const double *src, int32_t on_device);
/* The reference has no state accessor (plotfiles only, PlotExporter.hpp:54-59). */
tsp_status tsp_get_state(tsp_solver *s, double *dst, int32_t on_device);

/* Initial data from expressions, evaluated on the device: replaces
 * SingleLevelIntegrator::initialise_impl (SingleLevelIntegrator.hpp:75-93)
 * with a FunctionInit (init/FunctionInit.hpp:34-101), whose amrex::Parser
 * executors the reference calls on the device at every cell centre
 * x = prob_lo + (i + 0.5) dx. Each component is a postfix program of
 * tsp_expr_op (operands first, then the operation): NUM pushes `value`, X/Y/Z
 * push the cell centre's coordinate, binary codes pop b then a and push
 * a op b (POW: pow(a, b), ATAN2: atan2(a, b), MIN/MAX: fmin/fmax), unary
 * codes replace the top value (HEAVISIDE: 1 if v > 0 else 0). */
typedef enum tsp_expr_code {
    TSP_EXPR_NUM = 0, TSP_EXPR_X = 1, TSP_EXPR_Y = 2, TSP_EXPR_Z = 3,
    TSP_EXPR_ADD = 4, TSP_EXPR_SUB = 5, TSP_EXPR_MUL = 6, TSP_EXPR_DIV = 7, TSP_EXPR_POW = 8,
    TSP_EXPR_ATAN2 = 9, TSP_EXPR_MIN = 10, TSP_EXPR_MAX = 11,
    TSP_EXPR_NEG = 12, TSP_EXPR_SQRT = 13, TSP_EXPR_EXP = 14, TSP_EXPR_LOG = 15, TSP_EXPR_LOG10 = 16,
    TSP_EXPR_SIN = 17, TSP_EXPR_COS = 18, TSP_EXPR_TAN = 19, TSP_EXPR_ASIN = 20, TSP_EXPR_ACOS = 21,
    TSP_EXPR_ATAN = 22, TSP_EXPR_SINH = 23, TSP_EXPR_COSH = 24, TSP_EXPR_TANH = 25, TSP_EXPR_ABS = 26,
    TSP_EXPR_FLOOR = 27, TSP_EXPR_CEIL = 28, TSP_EXPR_HEAVISIDE = 29
} tsp_expr_code;
typedef struct tsp_expr_op {
    int32_t code;     /* tsp_expr_code */
    int32_t reserved; /* 0 */
    double value;     /* TSP_EXPR_NUM only */
} tsp_expr_op;
#define TSP_EXPR_MAX_STACK 16  /* values a program may hold at once */
#define TSP_EXPR_MAX_OPS 65536 /* operations of all components together */
/* Component c's program is ops[comp_start[c] .. comp_start[c+1]), comp_start[0] = 0.
 * index_lo: the domain's first cell index (amrex::Box smallEnd; NULL: 0, 0, 0).
 * The programs are checked first (TSP_INVALID_ARG: unknown code, underflow,
 * not exactly one value left; TSP_UNSUPPORTED: deeper than TSP_EXPR_MAX_STACK).
 * Writes state_old's valid cells and resets time, step count and register roles
 * like tsp_set_state; returns when the values are written. */
tsp_status tsp_init_expr(tsp_solver *s, const tsp_expr_op *ops, const int32_t *comp_start, const int32_t index_lo[3]);

/* Replaces SingleLevelIntegrator::FillPatch (SingleLevelIntegrator.hpp:153-163)
 * on the register with role `role` (tsp_role): periodic images then the
 * physical-boundary functor, faces, edges and corners (a TSP_BC_CUSTOM
 * functor at the solver's current time). */
tsp_status tsp_fill_ghosts(tsp_solver *s, int32_t role);

/* Replaces the physical-boundary step of FillPatch for a boundary functor the
 * library does not carry (TSP_BC_CUSTOM): the amrex::PhysBCFunct over
 * amrex::GpuBndryFuncFab<BCType>(model.get_bc()) that SingleLevelIntegrator.hpp:153-163
 * builds. After each fill's periodic step the library calls `fn` on its
 * stream with the register just filled and the fill's time: `time` before a
 * step (pre_step, SingleLevelIntegrator.hpp:102-108), time + dt * a_fraction[s]
 * before RK4 stage s >= 1 (integrators/SingleLevelRK4Integrator.hpp:119-125).
 * The callback launches the caller's kernel, which applies the functor to
 * every ghost cell outside the domain along a non-periodic axis (faces,
 * edges and corners); those cells hold, on entry, what the register last
 * held there. `fn` returns 0 on success; any other value stops the step with
 * TSP_HIP_ERROR. A TSP_BC_CUSTOM solver steps only after a callback is set
 * (TSP_INVALID_ARG otherwise), reads ghost memory (ghost_mode 1; stage pairs
 * and ghost_mode 0 are refused), steps eagerly (graph replay, which would
 * freeze the fill times, returns TSP_UNSUPPORTED), and cannot join a batch,
 * a level or an adjoint solve. */
typedef struct tsp_bc_view {
    int32_t n[3];        /* valid cells per axis */
    int32_t ghost;       /* ghost width G */
    int32_t ncomp;
    int32_t periodic[3];
    double time;         /* the fill's time */
    double *reg;         /* the register filled: element (i,j,k,c), ghosts included
                            (-G <= i < n + G), at reg[i + j*sj + k*sk + c*sc] */
    int64_t sj, sk, sc;
} tsp_bc_view;
typedef int (*tsp_bc_fn)(void *user, const tsp_bc_view *view, void *hip_stream);
tsp_status tsp_set_bc_callback(tsp_solver *s, tsp_bc_fn fn, void *user);
/* Copy a whole ghosted register (or its valid region) to host, for tests. */
tsp_status tsp_get_register(tsp_solver *s, int32_t role, double *dst_valid_plus_ghost);

/* Replaces `nsteps` calls of Integrator::step() (Integrator.hpp:91-106) on
 * SingleLevelRK4Integrator::computeNewState (SingleLevelRK4Integrator.hpp:95-157). */
tsp_status tsp_rk4_steps(tsp_solver *s, int64_t nsteps);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). `nsteps` RK4 steps of `state_in` into `state_out` (both
 * device pointers, the valid layout of tsp_set_state) with the solver's model,
 * kernels and configuration, enqueued on the solver's stream; bit-identical
 * to tsp_set_state + tsp_rk4_steps + tsp_get_state. The four RK4 registers
 * live in a file-scope device array of TSP_NOALLOC_SCRATCH_BYTES per device
 * instead of the solver's: the call allocates nothing, and a solver whose
 * registers need more is rejected with TSP_INVALID_ARG. The solver's own
 * registers, roles and clock are untouched. Every call on one device shares
 * that device's array (it keeps the last call's contents and is cleared at
 * the start of each call), so calls that overlap on one device race: order
 * them on one stream or synchronise between them. The current device must be
 * the one the solver was created on. */
#define TSP_NOALLOC_SCRATCH_BYTES (128ull << 20)
tsp_status tsp_rk4_steps_noalloc(tsp_solver *s, const double *state_in, double *state_out, int64_t nsteps);
/* Same on SingleLevelEulerIntegrator::computeNewState (SingleLevelEulerIntegrator.hpp:66-88). */
tsp_status tsp_euler_steps(tsp_solver *s, int64_t nsteps);

/* Integrator::time / Integrator::steps (Integrator.hpp:147-148). */
tsp_status tsp_get_time(tsp_solver *s, double *time, int64_t *steps);

tsp_status tsp_get_kernel(tsp_solver *s, tsp_kernel_cfg *cfg);
tsp_status tsp_set_kernel(tsp_solver *s, const tsp_kernel_cfg *cfg);

/* Time candidate stage-kernel configurations on this solver's shape (the state
 * is saved and restored) and keep the fastest; results are cached per
 * (device arch, system, stencil width, KO order, shape, ghost mode) in the
 * file named by TSP_TUNE_CACHE, if set. `steps` RK4 steps are timed per
 * candidate. The chosen configuration is written to *chosen if non-NULL. */
tsp_status tsp_autotune(tsp_solver *s, int32_t steps, tsp_kernel_cfg *chosen);

/* hipGraph replay of RK4 steps (no reference counterpart). The register
 * rotation of the reference has a period of two steps, so one graph holds two
 * steps (captured on first use for the current register roles);
 * tsp_graph_rk4_steps replays it and runs an odd remainder eagerly. Results
 * are bit-identical to tsp_rk4_steps. */
tsp_status tsp_graph_rk4_steps(tsp_solver *s, int64_t nsteps);
/* Capture and instantiate, without launching, the two-step graphs for the
 * current register roles and for those one RK4 step later, so that a later
 * tsp_graph_rk4_steps starting at either step parity only replays. */
tsp_status tsp_graph_prepare(tsp_solver *s);

/* Run `nsteps` eager RK4 steps recording HIP events around every stage
 * launch on the solver's stream; writes the summed device time of each stage
 * kernel (ms, 4 entries) and of any ghost-fill launches (ms, 1 entry). */
tsp_status tsp_profile_rk4(tsp_solver *s, int64_t nsteps, double *stage_ms, double *fill_ms);

/* Run `nsteps` RK4 steps (graph replay if use_graph != 0, else eager) with one
 * pair of HIP events recorded on the solver's stream around them; waits for
 * the stop event and writes the elapsed device time in ms. */
tsp_status tsp_time_rk4(tsp_solver *s, int64_t nsteps, int32_t use_graph, double *ms);

/* Block until the solver's stream is idle. */
tsp_status tsp_sync(tsp_solver *s);

/* The hipStream_t the solver enqueues its work on. */
tsp_status tsp_get_stream(tsp_solver *s, void **stream);

/* Device-side scope timer behind treesap::ScopedTimer
 * (util/ScopedTimer.hpp:34-149, which reads the host clock): a pair of HIP
 * events on one stream (NULL: the null stream). tsp_timer_start records the
 * first; tsp_timer_stop records the second, waits for it and writes the
 * device time between the two in ms, i.e. the time the stream spent on the
 * work enqueued between the calls, without the host's launch overhead or any
 * earlier work still queued when the scope began. A timer may be started
 * again after it stopped. */
typedef struct tsp_timer tsp_timer;
tsp_status tsp_timer_create(void *stream, tsp_timer **out);
tsp_status tsp_timer_start(tsp_timer *t);
tsp_status tsp_timer_stop(tsp_timer *t, double *elapsed_ms);
tsp_status tsp_timer_destroy(tsp_timer *t);

/* Device pointer of the register with the given role and its layout: element
 * (i,j,k,c), ghosts included (-G <= i < n+G), is at
 *   ptr[i + j*sj + k*sk + c*sc]   with ptr pointing at (0,0,0,0). */
tsp_status tsp_register_view(tsp_solver *s, int32_t role, double **ptr, int64_t *sj, int64_t *sk, int64_t *sc);

/* Batched mode (no reference counterpart): many independent models of the
 * (phi, Pi), (phi, u) and NonLinear systems, of uneven sizes, stencil widths
 * and KO orders, advanced in lockstep. Per RK4 stage there is one persistent
 * launch per (system, stencil width, KO order) group present in the batch,
 * covering every instance of the group; steps can
 * be replayed from a captured hipGraph. Each instance's result is
 * bit-identical to solving it alone with tsp_create/tsp_rk4_steps (the same
 * tile body runs). */
typedef struct tsp_batch tsp_batch;

typedef struct tsp_batch_cfg {
    int32_t tile_y;      /* 8 or 16 rows per workgroup (tile_x 64) */
    int32_t zchunk;      /* z planes per work item; 0 = min(nz, 32) per instance */
    int32_t grid_blocks; /* persistent grid size; 0 = one workgroup per work item */
    int32_t zalt;        /* 1: odd z chunks march downwards; 2: as 1, and forward RK4 stages 1 and 3
                            reverse every direction (as tsp_kernel_cfg.zalt) */
    int32_t halo_ring;   /* 1: halo-ring tile body where its LDS fits (as tsp_kernel_cfg.halo_ring) */
    int32_t narrow_tiles; /* 1: an x remainder of <= 32 cells runs as 32 x 2*tile_y tiles (no idle lanes) */
} tsp_batch_cfg;

/* Every descriptor must pass tsp_create's validation, use TSP_SYS_WAVE_PHIPI
 * and a (stencil_n, ko_order) pair in {(1,0),(1,1),(2,0),(2,2),(2,3),(3,0),
 * (3,3),(4,0),(4,4)}, with registers under 4 GiB each. All registers live in
 * one pool allocation. */
tsp_status tsp_batch_create(const tsp_desc *descs, int32_t count, void *stream, tsp_batch **out);
tsp_status tsp_batch_destroy(tsp_batch *b);
tsp_status tsp_batch_size(tsp_batch *b, int32_t *count, int64_t *cells);
/* Per-instance initial condition / state, same layout as tsp_set_state. */
tsp_status tsp_batch_set_state(tsp_batch *b, int32_t index, const double *src, int32_t on_device);
tsp_status tsp_batch_get_state(tsp_batch *b, int32_t index, double *dst, int32_t on_device);
tsp_status tsp_batch_get_time(tsp_batch *b, int32_t index, double *time, int64_t *steps);
/* Kernel configuration: every (system, stencil width, KO order) group has its
 * own (its launches and work items follow it). tsp_batch_set_kernel gives all
 * groups one configuration; tsp_batch_get_kernel reports the last one set
 * that way, or after tsp_batch_autotune the largest group's. The group
 * variants address one group, by system and (stencil_n, ko_order);
 * TSP_INVALID_ARG for a pair the batched kernels do not carry. Results never
 * depend on the configuration. */
tsp_status tsp_batch_get_kernel(tsp_batch *b, tsp_batch_cfg *cfg);
tsp_status tsp_batch_set_kernel(tsp_batch *b, const tsp_batch_cfg *cfg);
tsp_status tsp_batch_get_group_kernel(tsp_batch *b, int32_t system, int32_t stencil_n, int32_t ko_order,
                                      tsp_batch_cfg *cfg);
tsp_status tsp_batch_set_group_kernel(tsp_batch *b, int32_t system, int32_t stencil_n, int32_t ko_order,
                                      const tsp_batch_cfg *cfg);
/* `nsteps` RK4 steps of every instance: eager launches, or graph replay (two
 * steps per graph, odd remainder eager). */
tsp_status tsp_batch_rk4_steps(tsp_batch *b, int64_t nsteps);
tsp_status tsp_batch_graph_rk4_steps(tsp_batch *b, int64_t nsteps);
/* Capture and instantiate, without launching, the two-step graphs of both step
 * parities for the current configuration. */
tsp_status tsp_batch_graph_prepare(tsp_batch *b);
/* `nsteps` forward Euler steps of every instance (SingleLevelEulerIntegrator.hpp:66-88). */
tsp_status tsp_batch_euler_steps(tsp_batch *b, int64_t nsteps);
/* Tune every group on its own, since its stencil width and KO order decide
 * its best tile shape and LDS blocking: candidate configurations (tile
 * height, z chunk, persistent grid, halo ring, narrow tiles) over `steps`
 * graph-replayed RK4 steps in which only that group's kernels run, the three
 * fastest again over three times as many (also with zalt 2), and the fastest
 * kept for the group. Every instance's state, the register roles and the
 * clock are restored. Results are cached per group and (device arch, batch or
 * level, instance count, cells, groups present) in the file named by
 * TSP_TUNE_CACHE, if set. The largest group's configuration is written to
 * *chosen if non-NULL. */
tsp_status tsp_batch_autotune(tsp_batch *b, int32_t steps, tsp_batch_cfg *chosen);

/* A level split into boxes: replaces the BoxArray that
 * SingleLevelIntegrator(model, block_size) builds (SingleLevelIntegrator.hpp:47-49,
 * 172-176: the domain chopped into boxes of at most block_size cells per axis)
 * with its FillPatch across boxes (SingleLevelIntegrator.hpp:153-163). Each axis
 * is cut into the fewest pieces of at most block[a] cells, as equal as possible;
 * the boxes are the instances of a tsp_batch (box b = bx + nbx*(by + nby*bz)),
 * so tsp_batch_rk4_steps / _graph_rk4_steps / _euler_steps / _set_kernel /
 * _time_rk4 / _get_time apply. Before every stage the source register's
 * ghosts are exchanged between the boxes (neighbours, periodic images, edges
 * and corners) and the Dirichlet values are written outside non-periodic
 * faces; the stage kernels then read ghost memory. The absorbing layer is
 * measured in domain coordinates. Results are bit-identical to a single box
 * of the whole domain for every block size. (phi, Pi), (phi, u) and NonLinear
 * systems. */
tsp_status tsp_level_create(const tsp_desc *d, const int32_t block[3], void *stream, tsp_batch **out);
/* Box b's index range, lo_hi[6*b + 0..2] = lo (x, y, z), [6*b + 3..5] = hi (inclusive). */
tsp_status tsp_level_boxes(tsp_batch *b, int32_t *lo_hi);
/* The whole domain's valid state, layout as tsp_set_state; setting it restarts
 * the clock (initialise()). */
tsp_status tsp_level_set_state(tsp_batch *b, const double *src, int32_t on_device);
tsp_status tsp_level_get_state(tsp_batch *b, double *dst, int32_t on_device);
/* tsp_init_expr on every box of a level (cell centres in domain coordinates). */
tsp_status tsp_level_init_expr(tsp_batch *b, const tsp_expr_op *ops, const int32_t *comp_start,
                               const int32_t index_lo[3]);
tsp_status tsp_batch_time_rk4(tsp_batch *b, int64_t nsteps, int32_t use_graph, double *ms);
tsp_status tsp_batch_sync(tsp_batch *b);
/* The hipStream_t the batch (or level) enqueues its work on. */
tsp_status tsp_batch_get_stream(tsp_batch *b, void **stream);

/* Discrete adjoint (reverse mode) of the RK4 stepping, no reference
 * counterpart. For the linear (phi, Pi) system and a loss
 *   J = sum_{n=1..nsteps} j_n(phi^n at the receivers)
 * the gradient dJ/du^0 is
 *   lambda = 0; for n = nsteps..1: lambda += dj_n/du^n; lambda = G^T lambda,
 * with G^T the transpose of one RK4 step applied by the reverse-stage
 * recursion mu_3 = b_3 lambda', nu_j = L^T mu_j,
 * mu_{j-1} = b_{j-1} lambda' + a_j nu_j, lambda = lambda' + sum_j nu_j
 * (each reverse stage is one fused stage-kernel launch on the transpose
 * operator, with zero ghost values on non-periodic faces).
 * For the NonLinear (phi, pi) system (examples/02-temporaries/NonLinearRHS.hpp:52-71,
 * stencil_n 1..4) G^T is the transpose of the step's Jacobian at the forward
 * state: each reverse stage j adds N(a_j)^T mu_j to nu_j, N the linearised
 * -gamma pi |grad phi|^2 at the step's stage argument a_j, which the backward
 * pass recomputes from the stored forward states (two more launches per
 * stage, adjoint_nl.hip).
 * An adjoint object is bound to one solver (borrowed; it must outlive it) and
 * a fixed receiver set and step count; `receivers` holds nrec (i, j, k) valid
 * cell indices. */
typedef struct tsp_adjoint tsp_adjoint;
tsp_status tsp_adjoint_create(tsp_solver *s, const int32_t *receivers, int32_t nrec, int32_t nsteps,
                              int32_t checkpoint_every, tsp_adjoint **out);
tsp_status tsp_adjoint_destroy(tsp_adjoint *a);
/* Forward pass from the solver's current state: nsteps RK4 steps (graph
 * replayed), phi recorded at the receivers after every step into the trace
 * [nsteps][nrec], and the state stored every checkpoint_every steps (and at
 * the end). */
tsp_status tsp_adjoint_forward(tsp_adjoint *a);
tsp_status tsp_adjoint_trace(tsp_adjoint *a, double *dst, int32_t on_device);
/* Stored state of checkpoint `slot` (slot m = step m*checkpoint_every; the
 * last slot is the final state), valid layout as tsp_get_state. */
tsp_status tsp_adjoint_checkpoint(tsp_adjoint *a, int32_t slot, double *dst, int32_t on_device);
/* Backward pass: dtrace[n-1][r] = dJ/dphi^n(receiver r); writes dJ/du^0 (all
 * components, valid layout) to grad. With recompute != 0 every segment's
 * forward states are recomputed from its checkpoint before it is reversed
 * (the solver is left at the end of the forward pass); the linear system's
 * transpose does not depend on them, so the default skips that work. The
 * NonLinear system always recomputes: each segment's states are kept (one
 * segment, checkpoint_every states, at a time) and every reverse step
 * recomputes its stage arguments from them. */
tsp_status tsp_adjoint_backward(tsp_adjoint *a, const double *dtrace, int32_t dtrace_on_device, double *grad,
                                int32_t grad_on_device, int32_t recompute);
/* One transposed RK4 step, lam_out = G^T lam_in (valid layout), for
 * dot-product checks; for the NonLinear system linearised at the solver's
 * current state. */
tsp_status tsp_adjoint_apply(tsp_adjoint *a, const double *lam_in, double *lam_out, int32_t on_device);

/* Discrete adjoint of a whole batch (no reference counterpart): the single-box
 * adjoint above for every instance at once, each launch covering all
 * instances (one receiver gather or inject, one reverse-stage launch per
 * variant group). Bound to one batch (borrowed; it must outlive it); the nrec
 * receivers (i, j, k) are shared by all instances and must lie inside every
 * instance's box. Traces are [instance][step][receiver]; gradients and dtrace
 * follow the same instance-major order, gradients in the valid layout of
 * tsp_batch_get_state concatenated over instances. Every instance's trace and
 * gradient are bit-identical to a single-box tsp_adjoint on that instance. */
typedef struct tsp_batch_adjoint tsp_batch_adjoint;
tsp_status tsp_batch_adjoint_create(tsp_batch *b, const int32_t *receivers, int32_t nrec, int32_t nsteps,
                                    int32_t checkpoint_every, tsp_batch_adjoint **out);
tsp_status tsp_batch_adjoint_destroy(tsp_batch_adjoint *a);
tsp_status tsp_batch_adjoint_forward(tsp_batch_adjoint *a);
tsp_status tsp_batch_adjoint_trace(tsp_batch_adjoint *a, double *dst, int32_t on_device);
tsp_status tsp_batch_adjoint_backward(tsp_batch_adjoint *a, const double *dtrace, int32_t dtrace_on_device,
                                      double *grad, int32_t grad_on_device, int32_t recompute);

/* Stepping of a model whose right-hand side the library does not carry
 * (TSP_SYS_CUSTOM): the library owns the registers, the FillPatch before every
 * stage (periodic images, then the Dirichlet/NoOp boundary values; without a
 * periodic axis and under NoOp/Dirichlet the ghost cells already hold those
 * values and the pass is skipped, so the stage kernel writes valid cells
 * only, as runStage does), the
 * register roles and the clock, exactly as tsp_rk4_steps / tsp_euler_steps do,
 * and calls `fn` once per stage to launch the caller's stage kernel on
 * `hip_stream` (SingleLevelRK4Integrator.hpp:110-157 runStage,
 * SingleLevelEulerIntegrator.hpp:66-88). The view holds everything that
 * kernel needs; element (i,j,k,c) of a register, ghosts included
 * (-G <= i < n + G), is at reg[i + j*sj + k*sk + c*sc]. `fn` returns 0 on
 * success; any other value stops the stepping with TSP_HIP_ERROR. */
typedef struct tsp_stage_view {
    int32_t stage;      /* 0..3: RK4 stage; 4: forward Euler */
    int32_t n[3];       /* valid cells per axis */
    int32_t ghost;      /* ghost width G */
    int32_t ncomp;
    double time;        /* time + dt * a_fraction[stage] (RK4), time (Euler) */
    double b;           /* dt * b_fraction[stage] (RK4), dt (Euler) */
    double a_next;      /* dt * a_fraction[stage + 1] (stages 0..2) */
    const double *src;  /* register the RHS reads, ghosts filled: init (stage 0, Euler) or arg */
    const double *init; /* state_old */
    double *acc;        /* state_new */
    double *argn;       /* arg_next (stages 0..2), else NULL */
    int64_t sj, sk, sc; /* element strides, the same for all registers */
} tsp_stage_view;
typedef int (*tsp_stage_fn)(void *user, const tsp_stage_view *view, void *hip_stream);
tsp_status tsp_rk4_steps_custom(tsp_solver *s, int64_t nsteps, tsp_stage_fn fn, void *user);
tsp_status tsp_euler_steps_custom(tsp_solver *s, int64_t nsteps, tsp_stage_fn fn, void *user);

/* Algorithmic HBM bytes of one RK4 step per valid cell (128 * ncomp). */
int64_t tsp_rk4_bytes_per_cell(const tsp_desc *d);

#ifdef __cplusplus
}
#endif

#endif /* TREESAP_HIP_H */
