// batch_kernel.hpp -- one launch per RK4 stage for many independent boxes.
//
// Every instance is a separate model of the (phi, Pi), (phi, u) or NonLinear
// (phi, pi) system with its own box size, stencil width, KO order and
// strength, boundary handling and time step; all step in
// lockstep, so the register roles (state_old, state_new, arg, arg_next) rotate
// identically for all of them and travel as kernel arguments. The work list
// holds (instance, tile, z chunk) items, largest instances first; a grid of
// `gridDim.x` persistent blocks strides over it, dealing runs of consecutive
// items to one XCD (xcd_item). Each item runs exactly the
// tile body of the single-box kernel (FusedStage::run_tile), so a batched
// solve is bit-identical to solving every instance on its own.
#pragma once

#include "stage_kernel.hpp"

namespace tsp {

struct BatchInstance {
    double *reg[4];          // register origins (element (0,0,0,0))
    const double *sigma_tab; // ABL rate table or null
    int32_t nx, ny, nz, G, sj, sk, sc;
    int32_t variant;         // work-list group: batch_group(system, stencil width, KO order)
    int32_t periodic[3];
    int32_t dirichlet, abl_T, zchunk;
    int32_t off[3], dom[3]; // a box of a level: its offset in the domain and the domain's cells per axis
    double dt;
    double gf1[3], gf2[3];
    double ko_strength, gamma;
    double target[TSP_MAX_COMP];
};

struct BatchWork {
    int32_t inst, x0, y0, zc;
    int32_t narrow; // 1: a 32 x 2*TY tile (the 32-wide x remainder of a box)
};

struct BatchArgs {
    const BatchInstance *inst;
    const BatchWork *work;
    int32_t nwork;
    int32_t stage;           // StageKind
    int32_t src, init, acc, argn; // register indices of the roles for this stage
    int32_t zalt;
    int32_t ghost_mode;
};

// (stencil width, KO order) pairs the batched kernel carries
constexpr int kBatchVariants = 9;
constexpr int kBatchN[kBatchVariants] = {1, 1, 2, 2, 2, 3, 3, 4, 4};
constexpr int kBatchR[kBatchVariants] = {0, 1, 0, 2, 3, 0, 3, 0, 4};

inline int batch_variant(int n, int r) {
    for (int v = 0; v < kBatchVariants; ++v)
        if (kBatchN[v] == n && kBatchR[v] == r) return v;
    return -1;
}

// systems the batched kernels carry; a work-list group is one (system, variant)
// pair, group = slot * kBatchVariants + variant with the (phi, Pi) system at slot 0
constexpr int kBatchSystems = 3;
constexpr int kBatchGroups = kBatchSystems * kBatchVariants;
inline int batch_system_slot(int system) {
    return system == TSP_SYS_WAVE_PHIPI ? 0 : system == TSP_SYS_WAVE_PHIU ? 1 : system == TSP_SYS_NONLINEAR ? 2 : -1;
}
inline int batch_group(int system, int n, int r) {
    const int s = batch_system_slot(system), v = batch_variant(n, r);
    return s < 0 || v < 0 ? -1 : s * kBatchVariants + v;
}

// Work item of block slot w. The hardware deals blocks to the 8 XCDs round
// robin (slot w runs on XCD w mod 8); items go to the XCDs in runs of
// XCD_RUN consecutive items, so neighbouring tiles of one z chunk of an
// instance share one XCD's L2 (their halo rows are each other's queue heads)
// while every XCD still gets a share of the large and the small instances.
// A permutation of [0, nwork); slots past the last full round keep their item.
constexpr int XCD_RUN = 16;
__device__ __forceinline__ int xcd_item(int w, int nwork) {
    const int full = nwork / (8 * XCD_RUN) * (8 * XCD_RUN);
    if (w >= full) return w;
    const int x = w % 8, r = w / 8; // XCD of the slot and its rank among that XCD's slots
    return (r / XCD_RUN) * (8 * XCD_RUN) + x * XCD_RUN + r % XCD_RUN;
}

// One (stencil width, KO order) group: a persistent grid strides over the
// group's work items. RS > 0: the halo-ring tile body. Items are 64 x TY
// tiles or, flagged narrow, 32 x 2*TY tiles (same thread count).
// SM: compile-time stage kind (forward stage 0 or 3) or -1 (from the arguments).
// ADJ: reverse stage j = b.stage of the discrete adjoint (Sys the transposed
// system), with the coefficients the single-box adjoint forms from the
// instance's dt.
template <class Sys, int N, int R, int TY, int RS, int SM = -1, bool ADJ = false> struct BatchedStage {
    using FS = FusedStage<Sys, N, R, TY, ADJ, RS, true, 64, SM>;
    using NFS = FusedStage<Sys, N, R, 2 * TY, ADJ, RS, true, 32, SM>;
    static constexpr int LDS_TOTAL = FS::LDS_TOTAL > NFS::LDS_TOTAL ? FS::LDS_TOTAL : NFS::LDS_TOTAL;
    static constexpr bool FITS = (size_t)LDS_TOTAL * sizeof(double) <= 144 * 1024;

    __device__ __forceinline__ static void run(const BatchInstance *__restrict__ inst,
                                               const BatchWork *__restrict__ work, const BatchArgs &b, double *lds,
                                               double *sig_lds) {
        for (int w = blockIdx.x; w < b.nwork; w += gridDim.x) {
            const BatchWork wk = work[xcd_item(w, b.nwork)];
            const BatchInstance &I = inst[wk.inst];
            StageArgs a;
            a.src = I.reg[b.src];
            a.init = I.reg[b.init];
            a.acc = I.reg[b.acc];
            a.argn = b.argn >= 0 ? I.reg[b.argn] : nullptr;
            a.sigma_tab = I.sigma_tab;
            a.nx = I.nx;
            a.ny = I.ny;
            a.nz = I.nz;
            a.G = I.G;
            a.sj = I.sj;
            a.sk = I.sk;
            a.sc = I.sc;
            a.stage = b.stage;
            a.ghost_mode = b.ghost_mode;
            a.periodic[0] = I.periodic[0];
            a.periodic[1] = I.periodic[1];
            a.periodic[2] = I.periodic[2];
            a.dirichlet = I.dirichlet;
            a.abl_T = I.abl_T;
            for (int d = 0; d < 3; ++d) {
                a.off[d] = I.off[d];
                a.dom[d] = I.dom[d];
            }
            a.zchunk = I.zchunk;
            a.tiles_x = a.tiles_y = a.nzc = 0;
            a.xcd_remap = 0;
            a.zalt = b.zalt;
            a.halo_ring = RS > 0;
            a.ptr_access = 0;
            // integrators/SingleLevelRK4Integrator.hpp:71-74,119-121 (the same
            // products the host forms for a single box)
            const int s = b.stage;
            if constexpr (ADJ) {
                // reverse stage j: lambda'-scale b_3 on the first, next argument
                // b_{j-1} lambda' + a_j nu (b = dt*{1/6,1/3,1/3,1/6}, a = dt*{0,1/2,1/2,1})
                a.b = 0.0;
                a.a_next = s > 0 ? I.dt * (s == 3 ? 1.0 : 0.5) : 0.0;
                a.src_scale = s == 3 ? I.dt * (1.0 / 6.0) : 1.0;
                a.c_init_argn = s > 0 ? I.dt * (s == 1 ? 1.0 / 6.0 : 1.0 / 3.0) : 0.0;
            } else {
                a.b = s == kEuler ? I.dt : I.dt * ((s == 0 || s == 3) ? 1.0 / 6.0 : 1.0 / 3.0);
                a.a_next = I.dt * (s == 2 ? 1.0 : 0.5);
                a.c_init_argn = 0.0;
                a.src_scale = 1.0;
            }
            a.gf1[0] = I.gf1[0];
            a.gf1[1] = I.gf1[1];
            a.gf1[2] = I.gf1[2];
            a.gf2[0] = I.gf2[0];
            a.gf2[1] = I.gf2[1];
            a.gf2[2] = I.gf2[2];
            a.ko_strength = I.ko_strength;
            a.gamma = I.gamma;
#pragma unroll
            for (int c = 0; c < TSP_MAX_COMP; ++c) a.target[c] = I.target[c];
            const int z0 = wk.zc * I.zchunk;
            const int z1 = min(z0 + I.zchunk, I.nz);
            __syncthreads(); // the previous item's waves may still read LDS
            const double *sig = FS::stage_sigma(I.sigma_tab, I.abl_T, sig_lds);
            // zalt 2: forward stages 1 and 3 also reverse every direction (as the single-box kernel)
            const bool flip = !ADJ && b.zalt == 2 && (b.stage == kStage1 || b.stage == kStage3);
            const bool down = (b.zalt != 0 && (wk.zc & 1)) != flip;
            if (wk.narrow) {
                if (down) NFS::template run_tile<true>(a, wk.x0, wk.y0, z0, z1, lds, sig);
                else NFS::template run_tile<false>(a, wk.x0, wk.y0, z0, z1, lds, sig);
            } else {
                if (down) FS::template run_tile<true>(a, wk.x0, wk.y0, z0, z1, lds, sig);
                else FS::template run_tile<false>(a, wk.x0, wk.y0, z0, z1, lds, sig);
            }
        }
    }
};

// The instance table and work list are separate restrict-qualified arguments:
// nothing the kernel stores can alias them, so their uniform loads can be
// scalar loads.
template <class Sys, int N, int R, int TY, int RS = 0, int SM = -1, bool ADJ = false>
__global__ __launch_bounds__(64 * TY) void batched_stage_kernel(const BatchInstance *__restrict__ inst,
                                                                const BatchWork *__restrict__ work,
                                                                const BatchArgs b) {
    __shared__ double lds[BatchedStage<Sys, N, R, TY, RS, SM, ADJ>::LDS_TOTAL];
    extern __shared__ double sig_lds[];
    BatchedStage<Sys, N, R, TY, RS, SM, ADJ>::run(inst, work, b, lds, sig_lds);
}

// one tile height: the halo-ring kernel when asked for and its LDS fits
template <class Sys, int N, int R, int TY, int SM, bool ADJ = false>
hipError_t launch_batched_tile(const BatchArgs &b, int grid, size_t shm, int halo_ring, hipStream_t stream) {
    constexpr int RS = ring_slots<TY, FusedStage<Sys, N, R, TY>::H>();
    using Ring = BatchedStage<Sys, N, R, TY, RS, SM, ADJ>;
    if constexpr (RS > 0 && Ring::FITS) {
        if (halo_ring && sizeof(double) * Ring::LDS_TOTAL + shm <= max_block_lds()) {
            hipLaunchKernelGGL((batched_stage_kernel<Sys, N, R, TY, RS, SM, ADJ>), dim3(grid), dim3(64 * TY), shm,
                               stream, b.inst, b.work, b);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((batched_stage_kernel<Sys, N, R, TY, 0, SM, ADJ>), dim3(grid), dim3(64 * TY), shm, stream,
                       b.inst, b.work, b);
    return hipGetLastError();
}

// `max_abl_T`: the thickest absorbing layer in the group (sizes the LDS rate table)
template <class Sys, int N, int R, int SM, bool ADJ = false>
hipError_t launch_batched(const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring, hipStream_t stream) {
    const size_t shm = sigma_lds_bytes(max_abl_T);
    switch (tile_y) {
    case 8: return launch_batched_tile<Sys, N, R, 8, SM, ADJ>(b, grid, shm, halo_ring, stream);
    case 16: return launch_batched_tile<Sys, N, R, 16, SM, ADJ>(b, grid, shm, halo_ring, stream);
    default: return hipErrorInvalidValue;
    }
}

// every (stencil width, KO order) variant of one system's batch for one stage
// body; instantiated in batch.hip ((phi, Pi), SM = -1), batch_s0.hip and
// batch_s3.hip ((phi, Pi) stage-specific bodies), batch_phiu.hip and
// batch_nonlinear.hip (generic body), which the build compiles in parallel
template <class Sys, int SM>
hipError_t launch_batched_mode(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                               hipStream_t stream) {
    hipError_t e = hipErrorInvalidValue;
    static_for<0, kBatchVariants>([&]<int V>() {
        if (variant == V)
            e = launch_batched<Sys, kBatchN[V], kBatchR[V], SM>(b, tile_y, grid, max_abl_T, halo_ring, stream);
    });
    return e;
}
template <int SM>
hipError_t launch_batched_phipi_mode(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T,
                                     int halo_ring, hipStream_t stream) {
    return launch_batched_mode<SysPhiPi, SM>(variant, b, tile_y, grid, max_abl_T, halo_ring, stream);
}
hipError_t launch_batched_phiu(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                               hipStream_t stream);
hipError_t launch_batched_nonlinear(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T,
                                    int halo_ring, hipStream_t stream);
hipError_t launch_batched_phipi_s0(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                                   hipStream_t stream);
hipError_t launch_batched_phipi_s3(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                                   hipStream_t stream);

// launch one variant group of the (phi, Pi) batch
hipError_t launch_batched_phipi(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                                hipStream_t stream);
// launch one work-list group (system and variant) of a batch
hipError_t launch_batched_group(int group, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                                hipStream_t stream);
// one reverse stage (b.stage = j) of the discrete adjoint for one variant group
// (transposed system on the adjoint registers; batch_adjoint.hip)
hipError_t launch_batched_phipi_adjoint(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T,
                                        int halo_ring, hipStream_t stream);

// Receiver traces of a batched adjoint: block i serves instance i, reading
// register `role`; trace/dtrace rows are [instance][row][receiver] and every
// instance keeps its own row counter (counter[i]), advanced by its block.
hipError_t launch_batch_gather_trace(const BatchInstance *inst, int count, int role, const int *rec_off, int nrec,
                                     double *trace, int nrows, int *counter, hipStream_t stream);
hipError_t launch_batch_inject_trace(const BatchInstance *inst, int count, int role, const int *rec_off, int nrec,
                                     const double *dtrace, int nrows, int *counter, hipStream_t stream);
hipError_t launch_batch_set_counter(int *counter, int count, int v, hipStream_t stream);

// A level split into boxes (tsp_level_*): the instances of a batch are its
// boxes, instance b = bx + nbox[0] * (by + nbox[1] * bz) (level.hip).
struct LevelArgs {
    const BatchInstance *inst;
    const int32_t *axis_box; // box index along axis a of every domain index: [dom0 | dom1 | dom2]
    int32_t dom[3], nbox[3], periodic[3];
    int32_t dirichlet, G, nc, nboxes;
    double target[TSP_MAX_COMP];
};
// FillPatch of register `role` of every box: ghost cells from the owning box's
// valid cells (periodic images along periodic axes), Dirichlet values outside
// non-periodic faces; `max_shell` is the ghost-cell count of the largest box
hipError_t launch_level_exchange(const LevelArgs &L, int role, long long max_shell, hipStream_t stream);
// the domain's dense valid state into (scatter) or out of the boxes' register `role`
hipError_t launch_level_copy(const LevelArgs &L, int role, double *dense, bool scatter, hipStream_t stream);

} // namespace tsp
