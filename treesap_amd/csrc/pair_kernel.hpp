// pair_kernel.hpp -- two consecutive forward RK4 stages in one z march.
//
// The forward step of integrators/SingleLevelRK4Integrator.hpp:137-156 is
// four stages, each reading the previous stage's argument register with
// stencils. This kernel runs stages (0, 1) or (2, 3) of one step in one
// launch: a workgroup evaluates the first stage of the pair on a 32 x 32
// "extended" tile (its 32 - 2H x 32 - 2H output tile plus an H-wide ring
// that the neighbouring tiles also evaluate) and the second stage on the
// output tile only, reading the first stage's argument from registers (z)
// and LDS (x, y) instead of HBM. Per cell and step the pairs move 3 + 4
// register streams instead of the four stages' 3 + 5 + 5 + 3.
//
// Every cell value is the same expression of the same inputs as in the
// single-stage kernel (FusedStage), so results are bit-identical to it. Where
// the extended tile leaves the domain along a non-periodic axis, the first
// stage is not evaluated: its argument there is the intermediate register's
// ghost value, which is what the single-stage kernel of the second stage
// reads (ghost cells along non-periodic axes are never written by a stage);
// along periodic axes the wrapped cell is evaluated, as its owner does.
//
// Step t of a chunk [z0, z1) evaluates the first stage at plane
// ps = z0 - H + t (every extended cell) and the second stage at plane
// pc = ps - H - 1 (output cells), for t = 0 .. (z1 - z0) + 2H. Queues in
// registers, per cell:
//   q  : source planes ps-H .. ps+H of the first stage (head loaded a step ahead)
//   aq : first-stage argument planes pc-H .. pc+H, pushed as they are evaluated
// LDS: the source plane ps with its x/y halo, and the argument plane pc, each
// double buffered, so one barrier per step separates the writes of step t
// from the reads of step t-1; and a ring of H+1 output-tile planes with the
// first stage's acc values, each written and read back H+1 steps later by the
// same thread (no barrier involved).
//
// Register roles (the host passes them in StageArgs):
//   pair (0,1): src = state_old, init = the intermediate register (ghost reads
//               only), acc = state_new, argn = stage 1's arg_next
//   pair (2,3): src = stage 1's arg_next, init = state_old, acc = state_new,
//               argn = the intermediate register (ghost reads only)
#pragma once

#include "stage_kernel.hpp"

namespace tsp {

template <class Sys, int N, int R, int P> struct StagePair {
    static_assert(P == 0 || P == 2, "stage pairs (0, 1) and (2, 3)");
    static constexpr int H = N > R ? N : R;
    static constexpr int NC = Sys::NC;
    static_assert(Staging<Sys, R>::count() == NC, "every component is read with stencils (KO order > 0)");
    static constexpr int TE = 32;         // extended tile edge, one cell per thread and row
    static constexpr int TI = TE - 2 * H; // output tile edge
    // rows per thread: up to H = 3 the two queues of 2H+1 planes per cell fit
    // the 128 VGPRs of a 1024-thread block; at H = 4 they spill there, and
    // 512 threads with two rows each (rows r and r + 16) have 256 VGPRs
    static constexpr int RPT = H >= 4 ? 2 : 1;
    static constexpr int NT = TE * TE / RPT;
    static constexpr int RSTEP = TE / RPT; // row distance of a thread's rows
    static constexpr int LA = TE + 2 * H;  // source plane edge with its halo
    static constexpr int NHALO = 4 * H * TE;
    static constexpr int HPT = (NHALO + NT - 1) / NT;
    static constexpr int Q = 2 * H + 1;
    static constexpr int PA = NC * LA * LA; // doubles of one source plane buffer
    static constexpr int PB = NC * TE * TE; // doubles of one argument plane buffer
    static constexpr int PC = NC * TI * TI; // doubles of one acc plane of the ring
    static constexpr int LDS_DOUBLES = 2 * PA + 2 * PB + (H + 1) * PC;

    // cell view: z from a register queue, x/y from an LDS plane of row width W
    // and component stride CS (c0: the centre cell of component 0)
    template <int W, int CS> struct View {
        const double (&q)[NC][Q];
        const double *c0;
        const StageArgs &a;
        __device__ __forceinline__ double gf1(int ax) const { return a.gf1[ax]; }
        __device__ __forceinline__ double gf2(int ax) const { return a.gf2[ax]; }
        template <int C, int AX, int S> __device__ __forceinline__ double at() const {
            if constexpr (S == 0) return q[C][H];
            else if constexpr (AX == 2) return q[C][H + S];
            else if constexpr (AX == 0) return c0[C * CS + S];
            else return c0[C * CS + S * W];
        }
    };

    // k = F<C>(src): the decorated right-hand side, term for term as FusedStage
    // (ABLDecorator.hpp:102-108, KODecorator.hpp:58-70)
    template <int C, class V>
    __device__ __forceinline__ static double rhs(const V &v, const StageArgs &a, double sigma, int T) {
        double k = Sys::template user<N, C>(v, a);
        double ko_term = 0.0;
        ko_term += KO<R, 0, C>(v);
        ko_term += KO<R, 1, C>(v);
        ko_term += KO<R, 2, C>(v);
        k = k + a.ko_strength * ko_term;
        if (T > 0) k = k - sigma * (v.q[C][H] - a.target[C]);
        return k;
    }

    __device__ __forceinline__ static double abl_sigma(const double *sig, int T, int cxy2, int z, int nz) {
        const int dzv = z - min(max(z, T), nz - 1 - T);
        const int d2 = cxy2 + dzv * dzv;
        return d2 == 0 ? 0.0 : sig[d2];
    }

    __device__ __forceinline__ static void run_tile(const StageArgs &a, int x0, int y0, int z0, int z1, double *lds,
                                                    const double *sig) {
        const int tid = threadIdx.x;
        const int lane = tid & (TE - 1), row0 = tid / TE;
        const int wrapx = a.periodic[0], wrapy = a.periodic[1], wrapz = a.periodic[2];
        const BufAccess mem(a);
        auto zmap = [&](int k) { return map_axis(k, a.nz, wrapz, a.G); };
        using Col = decltype(mem.col(0));

        // this thread's extended cells (rows row0 + r*RSTEP) and their columns
        // (wrapped along periodic axes); output cells: the inner TI x TI inside the box
        const int ex = x0 - H + lane;
        const int mx = map_axis(ex, a.nx, wrapx, a.G);
        const bool x_inner = lane >= H && lane < TE - H;
        int ey[RPT], my[RPT], ii[RPT];
        bool xy_in[RPT], inner[RPT], active[RPT];
        Col qh[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int row = row0 + r * RSTEP;
            ey[r] = y0 - H + row;
            my[r] = map_axis(ey[r], a.ny, wrapy, a.G);
            xy_in[r] = mx >= 0 && mx < a.nx && my[r] >= 0 && my[r] < a.ny;
            qh[r] = mem.col(mx + my[r] * a.sj);
            inner[r] = x_inner && row >= H && row < TE - H;
            active[r] = inner[r] && ex < a.nx && ey[r] < a.ny;
            ii[r] = (row - H) * TI + lane - H; // index in an acc plane (output cells)
        }

        // halo of the source plane (rows above/below, columns left/right; no corners)
        Col hh[HPT];
        int hlds[HPT];
        bool hok[HPT];
#pragma unroll
        for (int m = 0; m < HPT; ++m) {
            const int h = tid + m * NT;
            int hx, hy;
            if (h < 2 * H * TE) {
                const int rr = h / TE;
                hy = rr < H ? rr - H : TE + (rr - H);
                hx = h % TE;
            } else {
                const int e = h - 2 * H * TE;
                const int mm = e % (2 * H);
                hy = e / (2 * H);
                hx = mm < H ? mm - H : TE + (mm - H);
            }
            hok[m] = h < NHALO;
            const int hcol = hok[m] ? map_axis(x0 - H + hx, a.nx, wrapx, a.G) +
                                          map_axis(y0 - H + hy, a.ny, wrapy, a.G) * a.sj
                                    : mx + my[0] * a.sj;
            hh[m] = mem.col(hcol);
            hlds[m] = (hy + H) * LA + hx + H;
        }

        // ABL: x/y part of the squared distance to the bulk box, for the
        // first-stage cells (wrapped) and the output cells
        const int T = a.abl_T;
        int sxy2[RPT], oxy2[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            sxy2[r] = 0;
            oxy2[r] = 0;
            if (T > 0) {
                const int dx1 = mx - min(max(mx, T), a.nx - 1 - T), dy1 = my[r] - min(max(my[r], T), a.ny - 1 - T);
                sxy2[r] = dx1 * dx1 + dy1 * dy1;
                const int dx2 = ex - min(max(ex, T), a.nx - 1 - T), dy2 = ey[r] - min(max(ey[r], T), a.ny - 1 - T);
                oxy2[r] = dx2 * dx2 + dy2 * dy2;
            }
        }

        double q[RPT][NC][Q], qn[RPT][NC], hv[NC][HPT], aq[RPT][NC][Q];
        double *const ring = lds + 2 * PA + 2 * PB;
        auto load_head = [&](int k) {
            const auto zq = mem.plane(zmap(k));
#pragma unroll
            for (int r = 0; r < RPT; ++r) static_for<0, NC>([&]<int C>() { qn[r][C] = mem.template src<C>(qh[r], zq); });
        };
        auto load_halo = [&](int k) {
            const auto zh = mem.plane(zmap(k));
            static_for<0, NC>([&]<int C>() {
#pragma unroll
                for (int m = 0; m < HPT; ++m) hv[C][m] = mem.template src<C>(hh[m], zh);
            });
        };

        // prologue: source planes z0-2H-1 .. z0-1 (the first is shifted out
        // unread), the head of step 0 (z0) and the halo of step 0 (z0-H)
#pragma unroll
        for (int r = 0; r < RPT; ++r)
            static_for<0, NC>([&]<int C>() {
                q[r][C][0] = 0.0;
#pragma unroll
                for (int m = 1; m < Q; ++m)
                    q[r][C][m] = mem.template src<C>(qh[r], mem.plane(zmap(z0 - 2 * H - 1 + m)));
#pragma unroll
                for (int m = 0; m < Q; ++m) aq[r][C][m] = 0.0;
            });
        load_head(z0);
        load_halo(z0 - H);

        const int nstep = (z1 - z0) + 2 * H + 1;
        int aslot = 0; // acc ring slot of step t: t mod (H+1), read for plane pc, then written for ps
        for (int t = 0; t < nstep; ++t) {
            double *const acc_plane = ring + aslot * PC;
            const int ps = z0 - H + t; // first stage
            const int pc = ps - H - 1; // second stage
            const bool last = t == nstep - 1;
            // q[.][.][0] before the shift is source plane pc (state_old there for pair (0,1))
            double ucen[RPT][NC];
#pragma unroll
            for (int r = 0; r < RPT; ++r)
                static_for<0, NC>([&]<int C>() {
                    ucen[r][C] = q[r][C][0];
#pragma unroll
                    for (int m = 0; m < Q - 1; ++m) q[r][C][m] = q[r][C][m + 1];
                    q[r][C][Q - 1] = qn[r][C];
                });
            const int zs = zmap(ps);
            const bool z_in = zs >= 0 && zs < a.nz;
            const auto zsp = mem.plane(zs);
            // centre values of the first stage: state_old (pair (2,3)) and acc
            // (pair (2,3), output cells), or the intermediate register's ghost value
            double iv[RPT][NC], av[RPT][NC], mv[RPT][NC];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                static_for<0, NC>([&]<int C>() {
                    iv[r][C] = 0.0;
                    av[r][C] = 0.0;
                    mv[r][C] = 0.0;
                });
                if (!last) {
                    if constexpr (P == 2) {
                        static_for<0, NC>([&]<int C>() { iv[r][C] = mem.template init<C>(qh[r], zsp); });
                        if (inner[r]) static_for<0, NC>([&]<int C>() { av[r][C] = mem.template acc<C>(qh[r], zsp); });
                    }
                    if (!(xy_in[r] && z_in)) {
                        static_for<0, NC>([&]<int C>() {
                            if constexpr (P == 0) mv[r][C] = mem.template init<C>(qh[r], zsp);
                            else mv[r][C] = BufAccess::ld<0>(mem.rn, qh[r], zsp + C * mem.scb);
                        });
                    }
                }
            }
            load_head(ps + 1 + H);
            double *A = lds + (t & 1) * PA;
            double *B = lds + 2 * PA + (t & 1) * PB;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int row = row0 + r * RSTEP;
                    A[c * LA * LA + (row + H) * LA + lane + H] = q[r][c][H];
                    B[c * TE * TE + row * TE + lane] = aq[r][c][H];
                }
#pragma unroll
                for (int m = 0; m < HPT; ++m)
                    if (hok[m]) A[c * LA * LA + hlds[m]] = hv[c][m];
            }
            load_halo(ps + 1);
            __syncthreads();

            // second stage at plane pc (inside the chunk from step 2H+1 on)
            if (t >= 2 * H + 1) {
                const auto zcp = mem.plane(pc);
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    if (!active[r]) continue;
                    const int row = row0 + r * RSTEP;
                    const View<TE, TE * TE> v{aq[r], B + row * TE + lane, a};
                    const double sigma = T > 0 ? abl_sigma(sig, T, oxy2[r], pc, a.nz) : 0.0;
                    static_for<0, NC>([&]<int C>() {
                        const double k = rhs<C>(v, a, sigma, T);
                        mem.template put_acc<C>(acc_plane[C * TI * TI + ii[r]] + a.b1 * k, qh[r], zcp);
                        if constexpr (P == 0) mem.template put_argn<C>(ucen[r][C] + a.a_next1 * k, qh[r], zcp);
                    });
                    // one row's stencil operands live at a time
                    __builtin_amdgcn_sched_barrier(0);
                }
            }

            // first stage at plane ps, pushed into the argument and acc queues
            if (!last) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int row = row0 + r * RSTEP;
                    double argv[NC], accv[NC];
                    if (xy_in[r] && z_in) {
                        const View<LA, LA * LA> v{q[r], A + (row + H) * LA + lane + H, a};
                        const double sigma = T > 0 ? abl_sigma(sig, T, sxy2[r], zs, a.nz) : 0.0;
                        static_for<0, NC>([&]<int C>() {
                            const double k = rhs<C>(v, a, sigma, T);
                            const double initc = P == 0 ? q[r][C][H] : iv[r][C];
                            if constexpr (P == 0) accv[C] = initc + a.b * k;
                            else accv[C] = av[r][C] + a.b * k;
                            argv[C] = initc + a.a_next * k;
                        });
                    } else {
                        static_for<0, NC>([&]<int C>() {
                            argv[C] = mv[r][C];
                            accv[C] = 0.0;
                        });
                    }
                    static_for<0, NC>([&]<int C>() {
#pragma unroll
                        for (int m = 0; m < Q - 1; ++m) aq[r][C][m] = aq[r][C][m + 1];
                        aq[r][C][Q - 1] = argv[C];
                        if (inner[r]) acc_plane[C * TI * TI + ii[r]] = accv[C];
                    });
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            aslot = aslot == H ? 0 : aslot + 1;
        }
    }

    __device__ __forceinline__ static void run(const StageArgs &a) {
        __shared__ double lds[LDS_DOUBLES];
        extern __shared__ double sig_lds[];
        const int nb = gridDim.x;
        int bid = blockIdx.x;
        if (a.xcd_remap) {
            const int per = nb / 8, rem = nb % 8, x = bid % 8;
            bid = x * per + (x < rem ? x : rem) + bid / 8;
        }
        const int tix = bid % a.tiles_x;
        const int t2 = bid / a.tiles_x;
        const int tiy = t2 % a.tiles_y;
        const int zc = t2 / a.tiles_y;
        const int z0 = zc * a.zchunk;
        const int z1 = min(z0 + a.zchunk, a.nz);
        const double *sig = FusedStage<Sys, N, R, 8>::stage_sigma(a.sigma_tab, a.abl_T, sig_lds);
        run_tile(a, tix * TI, tiy * TI, z0, z1, lds, sig);
    }
};

template <class Sys, int N, int R, int P>
__global__ __launch_bounds__((StagePair<Sys, N, R, P>::NT)) void stage_pair_kernel(const StageArgs a) {
    StagePair<Sys, N, R, P>::run(a);
}

// Launch the pair starting at stage `first` (0 or 2). hipErrorNotSupported,
// without launching, where the pair kernel does not apply: registers of 4 GiB
// or more, periodic axes shorter than 2H (one wrap must reach every read), or
// LDS beyond the device's per-block limit.
template <class Sys, int N, int R>
hipError_t launch_pair(const StageArgs &a0, int first, hipStream_t stream) {
    using SP = StagePair<Sys, N, R, 0>;
    StageArgs a = a0;
    if (a.ptr_access || !regs_fit_buffer(a, Sys::NC) || a.ghost_mode != 0) return hipErrorNotSupported;
    const int n[3] = {a.nx, a.ny, a.nz};
    for (int ax = 0; ax < 3; ++ax)
        if (a.periodic[ax] && n[ax] < 2 * SP::H) return hipErrorNotSupported;
    const size_t shm = sigma_lds_bytes(a.abl_T);
    if (sizeof(double) * SP::LDS_DOUBLES + shm > max_block_lds()) return hipErrorNotSupported;
    a.tiles_x = (a.nx + SP::TI - 1) / SP::TI;
    a.tiles_y = (a.ny + SP::TI - 1) / SP::TI;
    a.nzc = (a.nz + a.zchunk - 1) / a.zchunk;
    const long long nblocks = (long long)a.tiles_x * a.tiles_y * a.nzc;
    if (nblocks <= 0 || nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (first == 0)
        hipLaunchKernelGGL((stage_pair_kernel<Sys, N, R, 0>), dim3((unsigned)nblocks), dim3(SP::NT), shm, stream, a);
    else if (first == 2)
        hipLaunchKernelGGL((stage_pair_kernel<Sys, N, R, 2>), dim3((unsigned)nblocks), dim3(SP::NT), shm, stream, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

} // namespace tsp
