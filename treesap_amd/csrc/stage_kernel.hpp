// stage_kernel.hpp -- fused RK4 stage / Euler kernel for systems whose RHS
// reads neighbours along the coordinate axes only (first and unmixed second
// derivatives plus Kreiss-Oliger dissipation).
//
// What one launch computes is integrators/SingleLevelRK4Integrator.hpp:137-156
// (or SingleLevelEulerIntegrator.hpp:66-88) over the whole box: per valid
// cell, components in order 0..ncomp-1, k = F<c>(src) with F the decorated RHS
// ABLDecorator<KODecorator<User>> (ABLDecorator.hpp:102-108,
// KODecorator.hpp:58-70), then the stage axpys.
//
// Layout on the GPU (gfx950, wave64):
//   * a workgroup owns a 64 x TY column tile in (x, y) (or 32 x 32, two rows
//     per wave) and marches a chunk of z planes; wave w of the group owns
//     x-row y0+w, lane l owns x0+l, so every global load and store of a wave
//     is one contiguous 512-byte row;
//   * z neighbours live in a per-lane register queue of 2H+1 planes (each
//     plane of the source register is loaded from HBM once per chunk, one
//     step ahead of use);
//   * x and y neighbours come from an LDS copy of the current plane with an
//     H-wide halo along x and y (corners are not needed for axis stencils);
//     the halo of the next plane (or, with the halo ring, of a plane RS steps
//     further) is prefetched into registers while the current plane is
//     computed;
//   * ghosts are either virtual (out-of-domain indices are mapped: periodic
//     wrap, Dirichlet constant) or read from ghost memory after a fill pass.
// The arithmetic is the reference's expression for expression (weights times
// differences, grid factor last, left-fold sums), so every tile shape gives
// bit-identical results.
#pragma once

#include "common.hpp"

#include <type_traits>

namespace tsp {

// ---------------------------------------------------------------------------
// one-dimensional stencils on a cell view (util/stencil_impl.hpp:206-268,
// util/stencil.hpp:77-104); V::at<C, AX, S>() is component C at offset S
// along axis AX from the centre cell.
// ---------------------------------------------------------------------------
template <int N, int AX, int C, class V> __device__ __forceinline__ double D1(const V &v) {
    double st = 0.0;
    static_for<1, N + 1>([&]<int S>() {
        constexpr double w = cx::w1(N, S);
        st += w * (v.template at<C, AX, S>() - v.template at<C, AX, -S>());
    });
    return v.gf1(AX) * st;
}

template <int N, int AX, int C, class V> __device__ __forceinline__ double D2(const V &v) {
    constexpr double w0 = cx::w2(N, 0);
    double st = w0 * v.template at<C, AX, 0>();
    static_for<1, N + 1>([&]<int S>() {
        constexpr double w = cx::w2(N, S);
        st += w * (v.template at<C, AX, S>() + v.template at<C, AX, -S>());
    });
    return v.gf2(AX) * st;
}

template <int R, int AX, int C, class V> __device__ __forceinline__ double KO(const V &v) {
    constexpr double c0 = static_cast<double>(cx::ko_coef(R, 0));
    constexpr double norm = cx::ko_norm(R);
    double result = c0 * v.template at<C, AX, 0>();
    static_for<1, R + 1>([&]<int S>() {
        constexpr double cs = static_cast<double>(cx::ko_coef(R, S));
        result += cs * (v.template at<C, AX, S>() + v.template at<C, AX, -S>());
    });
    return result * norm * v.gf1(AX);
}

// ---------------------------------------------------------------------------
// registered axis-stencil systems
// ---------------------------------------------------------------------------

// (phi, Pi) wave: phi_t = Pi, Pi_t = contract_a D_aa phi (tensor_operations.hpp:45-61)
struct SysPhiPi {
    static constexpr int NC = 2;
    static constexpr bool user_stencil(int c) { return c == 0; }
    template <int N, int C, class V> __device__ __forceinline__ static double user(const V &v, const StageArgs &) {
        if constexpr (C == 0) {
            return v.template at<1, 0, 0>();
        } else {
            double result = 0.0;
            result += D2<N, 0, 0>(v);
            result += D2<N, 1, 0>(v);
            result += D2<N, 2, 0>(v);
            return result;
        }
    }
};

// examples/01-wave-eqn/WaveEqnRHS.hpp:37-45: phi_t = contract_a D_a u_a, (u_a)_t = D_a phi
struct SysPhiU {
    static constexpr int NC = 4;
    static constexpr bool user_stencil(int) { return true; }
    template <int N, int C, class V> __device__ __forceinline__ static double user(const V &v, const StageArgs &) {
        if constexpr (C == 0) {
            double result = 0.0;
            result += D1<N, 0, 1>(v);
            result += D1<N, 1, 2>(v);
            result += D1<N, 2, 3>(v);
            return result;
        } else {
            return D1<N, C - 1, 0>(v);
        }
    }
};

// examples/02-temporaries/NonLinearRHS.hpp:39-71
struct SysNonLinear {
    static constexpr int NC = 2;
    static constexpr bool user_stencil(int c) { return c == 0; }
    template <int N, int C, class V> __device__ __forceinline__ static double user(const V &v, const StageArgs &a) {
        if constexpr (C == 0) {
            return v.template at<1, 0, 0>(); // workspace.pi = pi.load(idx)
        } else {
            const double g0 = D1<N, 0, 0>(v);
            const double g1 = D1<N, 1, 0>(v);
            const double g2 = D1<N, 2, 0>(v);
            double grad_sq = 0.0;
            grad_sq += g0 * g0;
            grad_sq += g1 * g1;
            grad_sq += g2 * g2;
            double laplacian = 0.0;
            laplacian += D2<N, 0, 0>(v);
            laplacian += D2<N, 1, 0>(v);
            laplacian += D2<N, 2, 0>(v);
            return (laplacian - a.gamma * v.template at<1, 0, 0>() * grad_sq);
        }
    }
};

// Transpose of the linear (phi, Pi) operator, for the discrete adjoint. With
// homogeneous ghosts (zero Dirichlet values or periodic images) D_aa and the
// KO operator are symmetric and the ABL term is diagonal, so
//   [[KO - s, I], [Lap, KO - s]]^T = [[KO - s, Lap], [I, KO - s]]:
//   phibar_t = Lap(Pibar) (+ KO, ABL),  Pibar_t = phibar (+ KO, ABL).
struct SysPhiPiT {
    static constexpr int NC = 2;
    static constexpr bool user_stencil(int c) { return c == 1; }
    template <int N, int C, class V> __device__ __forceinline__ static double user(const V &v, const StageArgs &) {
        if constexpr (C == 0) {
            double result = 0.0;
            result += D2<N, 0, 1>(v);
            result += D2<N, 1, 1>(v);
            result += D2<N, 2, 1>(v);
            return result;
        } else {
            return v.template at<0, 0, 0>();
        }
    }
};

// ---------------------------------------------------------------------------
// kernel
// ---------------------------------------------------------------------------
template <class Sys, int R> struct Staging {
    static constexpr bool staged(int c) { return R > 0 || Sys::user_stencil(c); }
    static constexpr int count() {
        int n = 0;
        for (int c = 0; c < Sys::NC; ++c) n += staged(c) ? 1 : 0;
        return n;
    }
    static constexpr int slot(int c) {
        int n = 0;
        for (int i = 0; i < c; ++i) n += staged(i) ? 1 : 0;
        return n;
    }
};

// Index map of one axis. Ghost cells along non-periodic axes hold their
// boundary values permanently (they are written when the solver is created
// and never touched by a stage), so in the virtual-ghost mode only periodic
// axes are remapped: an index within H of the domain wraps once (H <= n). In
// the fill-pass mode ghost memory is read as is. Indices that no valid cell
// consumes (more than H outside; partial tiles) are clamped into the
// allocation, their values are never used.
__device__ __forceinline__ int map_axis(int i, int n, int wrap, int G) {
    if (i < 0 || i >= n) {
        if (wrap) i += i < 0 ? n : -n;
        i = min(max(i, -G), n + G - 1);
    }
    return i;
}

// Register access. Elements are addressed as (column, plane, component):
// column = i + j*sj (i, j >= -G), plane index k >= -G.
//
// BufAccess: one buffer descriptor per register, based at element (-G,-G,-G)
// of component 0; the lane's column becomes a loop-invariant byte offset in a
// VGPR and the plane and component a scalar byte offset, so an access costs no
// vector address arithmetic. Needs every register to span less than 4 GiB.
struct BufAccess {
    static constexpr int kNT = 2; // nontemporal cache policy of a buffer access
    __amdgpu_buffer_rsrc_t rs, ri, ra, rn;
    uint32_t cb, skb, scb;
    int G;
    __device__ __forceinline__ static __amdgpu_buffer_rsrc_t rsrc(const double *origin, const StageArgs &a) {
        const int64_t shift = (int64_t)a.G * (1 + (int64_t)a.sj + (int64_t)a.sk);
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(origin - shift), 0, -1, 0x00020000);
    }
    __device__ __forceinline__ explicit BufAccess(const StageArgs &a)
        : rs(rsrc(a.src, a)), ri(rsrc(a.init, a)), ra(rsrc(a.acc, a)), rn(rsrc(a.argn ? a.argn : a.acc, a)),
          cb((uint32_t)(a.G + a.G * a.sj)), skb((uint32_t)a.sk * 8u), scb((uint32_t)a.sc * 8u), G(a.G) {}
    __device__ __forceinline__ uint32_t col(int c) const { return ((uint32_t)c + cb) * 8u; }
    // the plane offset is uniform; readfirstlane pins it to an SGPR (a scalar
    // offset the compiler computed in VGPRs would need a waterfall loop)
    __device__ __forceinline__ uint32_t plane(int k) const {
        return __builtin_amdgcn_readfirstlane((uint32_t)(k + G) * skb);
    }
    template <int AUX> __device__ __forceinline__ static double ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, AUX));
    }
    __device__ __forceinline__ static void st(double v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, vo, so, kNT);
    }
    template <int C> __device__ __forceinline__ double src(uint32_t c, uint32_t p) const { return ld<0>(rs, c, p + C * scb); }
    template <int C> __device__ __forceinline__ double init(uint32_t c, uint32_t p) const { return ld<kNT>(ri, c, p + C * scb); }
    template <int C> __device__ __forceinline__ double acc(uint32_t c, uint32_t p) const { return ld<kNT>(ra, c, p + C * scb); }
    template <int C> __device__ __forceinline__ void put_acc(double v, uint32_t c, uint32_t p) const { st(v, ra, c, p + C * scb); }
    template <int C> __device__ __forceinline__ void put_argn(double v, uint32_t c, uint32_t p) const { st(v, rn, c, p + C * scb); }
};

// PtrAccess: plain 64-bit pointers, for registers of 4 GiB or more.
struct PtrAccess {
    const StageArgs &a;
    __device__ __forceinline__ explicit PtrAccess(const StageArgs &a_) : a(a_) {}
    __device__ __forceinline__ int col(int c) const { return c; }
    __device__ __forceinline__ int64_t plane(int k) const { return (int64_t)k * a.sk; }
    template <int C> __device__ __forceinline__ double src(int c, int64_t p) const { return (a.src + C * (int64_t)a.sc + p)[c]; }
    template <int C> __device__ __forceinline__ double init(int c, int64_t p) const {
        return __builtin_nontemporal_load(&(a.init + C * (int64_t)a.sc + p)[c]);
    }
    template <int C> __device__ __forceinline__ double acc(int c, int64_t p) const {
        return __builtin_nontemporal_load(&(a.acc + C * (int64_t)a.sc + p)[c]);
    }
    template <int C> __device__ __forceinline__ void put_acc(double v, int c, int64_t p) const {
        __builtin_nontemporal_store(v, &(a.acc + C * (int64_t)a.sc + p)[c]);
    }
    template <int C> __device__ __forceinline__ void put_argn(double v, int c, int64_t p) const {
        __builtin_nontemporal_store(v, &(a.argn + C * (int64_t)a.sc + p)[c]);
    }
};

// Whether every register of a launch spans less than 4 GiB (buffer access).
inline bool regs_fit_buffer(const StageArgs &a, int ncomp) {
    const int64_t span = (int64_t)a.G * (1 + (int64_t)a.sj + (int64_t)a.sk) + (int64_t)ncomp * a.sc;
    return span * (int64_t)sizeof(double) < (int64_t(1) << 32);
}

// RS > 0 (halo ring): the x/y halo of a plane is loaded RS+1 steps ahead of
// its use, close to when the neighbouring tiles load those cells as their own
// z-queue heads (H+1 steps ahead) and L2 still holds them, and written at
// once into that plane's buffer: the LDS holds a ring of RS+2 plane buffers
// (plane t in buffer t mod (RS+2)). A buffer is rewritten at the top of step
// t for plane t+RS; its previous plane, t-2, was last read during step t-2,
// which every wave has finished once any wave is past the barrier of step
// t-1. Without the ring (RS = 0) the two buffers alternate.
//
// TXW: tile width, 64 (one wave per x row) or 32 (each wave covers two rows;
// for boxes whose x extent leaves a 32-wide remainder).
//
// SM: the stage kind as a compile-time constant (kStage0, kStage3 of the
// forward RK4), or -1 for one body that reads it from the arguments. A fixed
// stage drops the loads and stores it does not need from the code, not just
// at run time (forward stage 0 +8 %, stage 3 +16 % on wave256).
template <class Sys, int N, int R, int TY, bool ADJ = false, int RS = 0, bool BUF = true, int TXW = 64, int SM = -1>
struct FusedStage {
    static_assert(TXW == 64 || TXW == 32, "tile width 64 or 32");
    static constexpr int TX = TXW;
    static constexpr int H = N > R ? N : R;
    static constexpr int NC = Sys::NC;
    static constexpr int NS = Staging<Sys, R>::count();
    static constexpr int LW = TX + 2 * H;
    static constexpr int LH = TY + 2 * H;
    static constexpr int NT = TX * TY;
    static constexpr int NHALO = 2 * H * (TX + TY);
    static constexpr int HPT = (NHALO + NT - 1) / NT;
    static constexpr int Q = 2 * H + 1;
    static constexpr int LDS_PLANE = NS * LH * LW;
    static constexpr int NSLOT = RS > 0 ? RS + 2 : 2; // plane buffers
    static constexpr int LDS_DOUBLES = NSLOT * LDS_PLANE;
    static constexpr int LDS_TOTAL = LDS_DOUBLES;
    // ring variants are built where the plane buffers stay within 144 KiB of LDS
    static constexpr bool FITS = (size_t)LDS_TOTAL * sizeof(double) <= 144 * 1024;

    using Plane = double[LH][LW];

    // Lookahead of the stage-specific forward bodies (run_tile2: two steps for
    // queue heads and halos, one for centre values); run_tile1 elsewhere.
    static constexpr int LA = (SM >= 0 && !ADJ) ? 2 : 1;

    // DOWN: the tile marches towards decreasing z; the queue then holds
    // q[m] = plane z + H - m instead of z - H + m.
    template <bool DOWN> struct View {
        const double (&q)[NS][Q];
        const Plane *lds;
        const double (&cen)[NC];
        const StageArgs &a;
        int lx, ly; // LDS coordinates of the centre cell
        __device__ __forceinline__ double gf1(int ax) const { return a.gf1[ax]; }
        __device__ __forceinline__ double gf2(int ax) const { return a.gf2[ax]; }
        template <int C, int AX, int S> __device__ __forceinline__ double at() const {
            if constexpr (S == 0) {
                return cen[C];
            } else {
                constexpr int sl = Staging<Sys, R>::slot(C);
                static_assert(Staging<Sys, R>::staged(C), "off-centre read of an unstaged component");
                if constexpr (AX == 2) return q[sl][DOWN ? H - S : H + S];
                else if constexpr (AX == 0) return lds[sl][ly][lx + S];
                else return lds[sl][ly + S][lx];
            }
        }
    };

    using Access = std::conditional_t<BUF, BufAccess, PtrAccess>;

    // Source value of component C at column handle `col` of plane handle `pl`
    // (the adjoint's first reverse stage reads b_3 * lambda' without
    // materialising it).
    template <int C, class Col, class Pl>
    __device__ __forceinline__ static double src_at(const StageArgs &a, const Access &m, Col col, Pl pl) {
        const double v = m.template src<C>(col, pl);
        if constexpr (ADJ) return a.src_scale * v;
        else return v;
    }

    // One TX x TY column tile over z planes [z0, z1). `lds_raw` holds
    // LDS_DOUBLES doubles of LDS; `sig` is the ABL rate table by squared
    // distance, resident in LDS (null without an ABL). Every thread of the
    // block must call this.
    //
    // Forward (RK4 stage s / Euler, integrators/SingleLevelRK4Integrator.hpp:142-154):
    //   k = F(src);  acc = (first ? init : acc) + b*k;  argn = init + a_next*k
    // Adjoint (reverse stage j of the discrete adjoint RK4, Sys = transpose system):
    //   nu = L^T mu;  acc = (first ? lambda' : acc) + nu;  argn = b_{j-1}*lambda' + a_j*nu
    //
    // Memory pipeline. While plane t is computed, the x/y halo of plane t+1
    // (t+RS+1 with the ring) and the z-queue head of plane t+H+1 are in
    // flight; the init/acc values of plane t are issued at the top of its step
    // and consumed at its end. The loads that carry values into the next step
    // are unconditional (their indices are clamped into the allocation; values
    // beyond the chunk are never used), so each one lands in a fixed register
    // and no copy forces an early wait. The memory counter completes in order,
    // so every wait targets a load issued a phase earlier and leaves the newer
    // ones in flight.
    template <bool DOWN>
    __device__ __forceinline__ static void run_tile(const StageArgs &a, int x0, int y0, int z0, int z1,
                                                    double *lds_raw, const double *sig) {
        if constexpr (LA == 2) run_tile2<DOWN>(a, x0, y0, z0, z1, lds_raw, sig);
        else run_tile1<DOWN>(a, x0, y0, z0, z1, lds_raw, sig);
    }

    template <bool DOWN>
    __device__ __forceinline__ static void run_tile1(const StageArgs &a, int x0, int y0, int z0, int z1,
                                                     double *lds_raw, const double *sig) {
        const int tid = threadIdx.x;
        const int lane = tid & (TX - 1);
        const int row = tid / TX;
        const int gi = x0 + lane, gj = y0 + row;
        const bool active = gi < a.nx && gj < a.ny;
        const int cnt = z1 - z0;
        // prefetches for steps past the chunk are never consumed: they reread
        // the last step's planes, which the caches hold, instead of new ones
        const int tlast = cnt - 1;

        const int stage = SM >= 0 ? SM : a.stage;
        bool is_first, write_argn, read_init;
        if constexpr (ADJ) {
            is_first = stage == kStage3;
            write_argn = stage > kStage0;
            read_init = write_argn || is_first;
        } else {
            is_first = stage == kStage0 || stage == kEuler;
            write_argn = stage < kStage3;
            read_init = write_argn && !is_first;
        }
        const bool read_acc = !is_first;
        const int wrapx = a.ghost_mode == 0 && a.periodic[0];
        const int wrapy = a.ghost_mode == 0 && a.periodic[1];
        const int wrapz = a.ghost_mode == 0 && a.periodic[2];
        const Access mem(a);
        auto zoff_of = [&](int k) { return mem.plane(map_axis(k, a.nz, wrapz, a.G)); };
        auto plane = [&](int t) { return DOWN ? z1 - 1 - t : z0 + t; };          // z of step t
        auto qplane = [&](int t) { return DOWN ? z1 - 1 - t - H : z0 + t + H; }; // queue head of step t

        // column offsets of this lane's own column and of its halo cells
        const int qcol = map_axis(gi, a.nx, wrapx, a.G) + map_axis(gj, a.ny, wrapy, a.G) * a.sj;
        const int ccol = active ? gi + gj * a.sj : qcol; // centre column (clamped for lanes outside the box)
        const auto qh = mem.col(qcol), ch = mem.col(ccol);
        decltype(mem.col(0)) hh[HPT];
        int hlds[HPT];
        bool hok[HPT];
#pragma unroll
        for (int m = 0; m < HPT; ++m) {
            const int h = tid + m * NT;
            int hx, hy;
            if (h < 2 * H * TX) { // rows above/below the tile
                const int r = h / TX;
                hy = r < H ? r - H : TY + (r - H);
                hx = h % TX;
            } else { // columns left/right of the tile
                const int e = h - 2 * H * TX;
                const int mm = e % (2 * H);
                hy = e / (2 * H);
                hx = mm < H ? mm - H : TX + (mm - H);
            }
            hok[m] = h < NHALO;
            const int hcol =
                hok[m] ? map_axis(x0 + hx, a.nx, wrapx, a.G) + map_axis(y0 + hy, a.ny, wrapy, a.G) * a.sj : qcol;
            hh[m] = mem.col(hcol);
            hlds[m] = (hy + H) * LW + hx + H;
        }

        // ABL damping: the x/y part of the squared distance to the bulk box is fixed per lane
        int cxy2 = 0;
        const int T = a.abl_T;
        if (T > 0) { // in domain coordinates (a box of a level sits at a.off)
            const int xg = gi + a.off[0], yg = gj + a.off[1];
            const int dxv = xg - min(max(xg, T), a.dom[0] - 1 - T);
            const int dyv = yg - min(max(yg, T), a.dom[1] - 1 - T);
            cxy2 = dxv * dxv + dyv * dyv;
        }

        double q[NS][Q], qn[NS], hv[NS][HPT];

        auto load_halo = [&](int t) {
            const auto zh = zoff_of(plane(t));
            static_for<0, NC>([&]<int C>() {
                if constexpr (Staging<Sys, R>::staged(C)) {
                    constexpr int sl = Staging<Sys, R>::slot(C);
#pragma unroll
                    for (int m = 0; m < HPT; ++m) hv[sl][m] = src_at<C>(a, mem, hh[m], zh);
                }
            });
        };
        auto load_queue = [&](int t) {
            const auto zq = zoff_of(qplane(t));
            static_for<0, NC>([&]<int C>() {
                if constexpr (Staging<Sys, R>::staged(C)) qn[Staging<Sys, R>::slot(C)] = src_at<C>(a, mem, qh, zq);
            });
        };

        // the halo of plane step p into its buffer
        auto put_halo = [&](int slot) {
            double *dst = lds_raw + slot * LDS_PLANE;
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int m = 0; m < HPT; ++m)
                    if (hok[m]) dst[s * LH * LW + hlds[m]] = hv[s][m];
        };

        // prologue: the queue as it stands before step 0 advances it (planes
        // -H-1..H-1 of the chunk; the first is shifted out unread) and the
        // queue head of step 0 in flight; the halo of step 0 in flight, or with
        // the ring, the halos of steps 0..RS-1 in their buffers and that of step RS in flight
        static_for<0, NC>([&]<int C>() {
            if constexpr (Staging<Sys, R>::staged(C)) {
                constexpr int sl = Staging<Sys, R>::slot(C);
                q[sl][0] = 0.0;
#pragma unroll
                for (int m = 1; m < Q; ++m)
                    q[sl][m] = src_at<C>(a, mem, qh, zoff_of(DOWN ? z1 + H - m : z0 - H - 1 + m));
            }
        });
        load_queue(0);
        if constexpr (RS > 0) {
#pragma unroll
            for (int p = 0; p < RS; ++p) {
                load_halo(min(p, tlast));
                put_halo(p);
            }
            load_halo(min(RS, tlast));
        } else {
            load_halo(0);
        }
        int cslot = 0;                // buffer of plane step t: t % NSLOT
        int hslot = RS % NSLOT;       // buffer of plane step t + RS

        for (int t = 0; t < cnt; ++t) {
            const int z = plane(t);
            const auto zp = mem.plane(z);
            // advance the queue to planes z-H..z+H with the head loaded during
            // step t-1. The shift comes first in the step so that the old head
            // is dead before the next head is loaded into the same registers
            // (a head loaded before the shift would need a copy that waits for
            // the load just issued).
#pragma unroll
            for (int s = 0; s < NS; ++s) {
#pragma unroll
                for (int m = 0; m < Q - 1; ++m) q[s][m] = q[s][m + 1];
                q[s][Q - 1] = qn[s];
            }
            // centre values of this plane: init, acc and the source of unstaged components
            double iv[NC], av[NC], cv[NC];
            static_for<0, NC>([&]<int C>() {
                if constexpr (!Staging<Sys, R>::staged(C)) cv[C] = src_at<C>(a, mem, ch, zp);
                if (read_init) iv[C] = mem.template init<C>(ch, zp);
                if (read_acc) av[C] = mem.template acc<C>(ch, zp);
            });
            load_queue(min(t + 1, tlast));
            // this plane's centre into its buffer, and the halo loaded during
            // step t-1 (of plane t+RS; of plane t without the ring) into the
            // buffer of its plane
            double *buf = lds_raw + cslot * LDS_PLANE;
#pragma unroll
            for (int s = 0; s < NS; ++s) buf[s * LH * LW + (row + H) * LW + lane + H] = q[s][H];
            put_halo(hslot);
            if constexpr (RS > 0) {
                load_halo(min(t + RS + 1, tlast));
            } else {
                load_halo(min(t + 1, tlast));
            }
            cslot = cslot == NSLOT - 1 ? 0 : cslot + 1;
            hslot = hslot == NSLOT - 1 ? 0 : hslot + 1;
            __syncthreads();

            if (active) {
                double cen[NC];
                static_for<0, NC>([&]<int C>() {
                    if constexpr (Staging<Sys, R>::staged(C)) cen[C] = q[Staging<Sys, R>::slot(C)][H];
                    else cen[C] = cv[C];
                });
                double sigma = 0.0;
                if (T > 0) {
                    const int zg = z + a.off[2];
                    const int dzv = zg - min(max(zg, T), a.dom[2] - 1 - T);
                    const int d2 = cxy2 + dzv * dzv;
                    sigma = d2 == 0 ? 0.0 : sig[d2];
                }
                const View<DOWN> v{q, reinterpret_cast<const Plane *>(buf), cen, a, lane + H, row + H};
                static_for<0, NC>([&]<int C>() {
                    double k = Sys::template user<N, C>(v, a);
                    if constexpr (R > 0) {
                        double ko_term = 0.0;
                        ko_term += KO<R, 0, C>(v);
                        ko_term += KO<R, 1, C>(v);
                        ko_term += KO<R, 2, C>(v);
                        k = k + a.ko_strength * ko_term;
                    }
                    if (T > 0) k = k - sigma * (cen[C] - a.target[C]);
                    const double initc = read_init ? iv[C] : cen[C];
                    double accv, argv = 0.0;
                    if constexpr (ADJ) {
                        accv = (is_first ? initc : av[C]) + k;
                        if (write_argn) argv = a.c_init_argn * initc + a.a_next * k;
                    } else {
                        if (is_first) accv = initc + a.b * k;
                        else accv = av[C] + a.b * k;
                        if (write_argn) argv = initc + a.a_next * k;
                    }
                    mem.template put_acc<C>(accv, ch, zp);
                    if (write_argn) mem.template put_argn<C>(argv, ch, zp);
                });
            }

        }
    }

    // run_tile with a two-step lookahead, for the stage-specific forward
    // bodies: the queue head and x/y halo a step consumes were requested two
    // steps earlier and its centre init/acc values one step earlier, in
    // alternating register sets (the plane loop is unrolled by two so that
    // every set has a fixed register name), so no wait in step t targets a
    // load issued in step t and twice as many bytes are in flight.
    template <bool DOWN>
    __device__ __forceinline__ static void run_tile2(const StageArgs &a, int x0, int y0, int z0, int z1,
                                                     double *lds_raw, const double *sig) {
        static_assert(SM >= 0 && !ADJ, "two-step lookahead is built for the forward stage-specific bodies");
        const int tid = threadIdx.x;
        const int lane = tid & (TX - 1);
        const int row = tid / TX;
        const int gi = x0 + lane, gj = y0 + row;
        const bool active = gi < a.nx && gj < a.ny;
        const int cnt = z1 - z0;
        const int tlast = cnt - 1;

        constexpr bool is_first = SM == kStage0 || SM == kEuler;
        constexpr bool write_argn = SM < kStage3;
        constexpr bool read_init = write_argn && !is_first;
        constexpr bool read_acc = !is_first;
        const int wrapx = a.ghost_mode == 0 && a.periodic[0];
        const int wrapy = a.ghost_mode == 0 && a.periodic[1];
        const int wrapz = a.ghost_mode == 0 && a.periodic[2];
        const Access mem(a);
        auto zoff_of = [&](int k) { return mem.plane(map_axis(k, a.nz, wrapz, a.G)); };
        auto plane = [&](int t) { return DOWN ? z1 - 1 - t : z0 + t; };
        auto qplane = [&](int t) { return DOWN ? z1 - 1 - t - H : z0 + t + H; };

        const int qcol = map_axis(gi, a.nx, wrapx, a.G) + map_axis(gj, a.ny, wrapy, a.G) * a.sj;
        const int ccol = active ? gi + gj * a.sj : qcol;
        const auto qh = mem.col(qcol), ch = mem.col(ccol);
        decltype(mem.col(0)) hh[HPT];
        int hlds[HPT];
        bool hok[HPT];
#pragma unroll
        for (int m = 0; m < HPT; ++m) {
            const int h = tid + m * NT;
            int hx, hy;
            if (h < 2 * H * TX) {
                const int r = h / TX;
                hy = r < H ? r - H : TY + (r - H);
                hx = h % TX;
            } else {
                const int e = h - 2 * H * TX;
                const int mm = e % (2 * H);
                hy = e / (2 * H);
                hx = mm < H ? mm - H : TX + (mm - H);
            }
            hok[m] = h < NHALO;
            const int hcol =
                hok[m] ? map_axis(x0 + hx, a.nx, wrapx, a.G) + map_axis(y0 + hy, a.ny, wrapy, a.G) * a.sj : qcol;
            hh[m] = mem.col(hcol);
            hlds[m] = (hy + H) * LW + hx + H;
        }

        int cxy2 = 0;
        const int T = a.abl_T;
        if (T > 0) { // in domain coordinates (a box of a level sits at a.off)
            const int xg = gi + a.off[0], yg = gj + a.off[1];
            const int dxv = xg - min(max(xg, T), a.dom[0] - 1 - T);
            const int dyv = yg - min(max(yg, T), a.dom[1] - 1 - T);
            cxy2 = dxv * dxv + dyv * dyv;
        }

        double q[NS][Q], qn[2][NS], hv[2][NS][HPT], iv[2][NC], av[2][NC], cv[2][NC];

        auto load_halo = [&](int t, double (&dst)[NS][HPT]) {
            const auto zh = zoff_of(plane(t));
            static_for<0, NC>([&]<int C>() {
                if constexpr (Staging<Sys, R>::staged(C)) {
                    constexpr int sl = Staging<Sys, R>::slot(C);
#pragma unroll
                    for (int m = 0; m < HPT; ++m) dst[sl][m] = src_at<C>(a, mem, hh[m], zh);
                }
            });
        };
        auto load_queue = [&](int t, double (&dst)[NS]) {
            const auto zq = zoff_of(qplane(t));
            static_for<0, NC>([&]<int C>() {
                if constexpr (Staging<Sys, R>::staged(C)) dst[Staging<Sys, R>::slot(C)] = src_at<C>(a, mem, qh, zq);
            });
        };
        auto load_centre = [&](int t, double (&ivd)[NC], double (&avd)[NC], double (&cvd)[NC]) {
            const auto zp = mem.plane(plane(t));
            static_for<0, NC>([&]<int C>() {
                if constexpr (!Staging<Sys, R>::staged(C)) cvd[C] = src_at<C>(a, mem, ch, zp);
                if constexpr (read_init) ivd[C] = mem.template init<C>(ch, zp);
                if constexpr (read_acc) avd[C] = mem.template acc<C>(ch, zp);
            });
        };
        auto put_halo = [&](int slot, const double (&src)[NS][HPT]) {
            double *dst = lds_raw + slot * LDS_PLANE;
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int m = 0; m < HPT; ++m)
                    if (hok[m]) dst[s * LH * LW + hlds[m]] = src[s][m];
        };

        // prologue: queue planes before step 0, heads of steps 0 and 1, centre
        // values of step 0; halos of steps 0 and 1 in flight, or with the ring
        // those of steps 0..RS-1 in their buffers and of steps RS, RS+1 in flight
        static_for<0, NC>([&]<int C>() {
            if constexpr (Staging<Sys, R>::staged(C)) {
                constexpr int sl = Staging<Sys, R>::slot(C);
                q[sl][0] = 0.0;
#pragma unroll
                for (int m = 1; m < Q; ++m)
                    q[sl][m] = src_at<C>(a, mem, qh, zoff_of(DOWN ? z1 + H - m : z0 - H - 1 + m));
            }
        });
        load_queue(0, qn[0]);
        load_queue(min(1, tlast), qn[1]);
        load_centre(0, iv[0], av[0], cv[0]);
        if constexpr (RS > 0) {
#pragma unroll
            for (int p = 0; p < RS; ++p) {
                load_halo(min(p, tlast), hv[0]);
                put_halo(p, hv[0]);
            }
            load_halo(min(RS, tlast), hv[0]);
            load_halo(min(RS + 1, tlast), hv[1]);
        } else {
            load_halo(0, hv[0]);
            load_halo(min(1, tlast), hv[1]);
        }
        int cslot = 0;
        int hslot = RS % NSLOT;

        auto step = [&]<int P>(int t) {
            const int z = plane(t);
            const auto zp = mem.plane(z);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
#pragma unroll
                for (int m = 0; m < Q - 1; ++m) q[s][m] = q[s][m + 1];
                q[s][Q - 1] = qn[P][s];
            }
            load_centre(min(t + 1, tlast), iv[1 - P], av[1 - P], cv[1 - P]);
            load_queue(min(t + 2, tlast), qn[P]);
            double *buf = lds_raw + cslot * LDS_PLANE;
#pragma unroll
            for (int s = 0; s < NS; ++s) buf[s * LH * LW + (row + H) * LW + lane + H] = q[s][H];
            put_halo(hslot, hv[P]);
            if constexpr (RS > 0) load_halo(min(t + RS + 2, tlast), hv[P]);
            else load_halo(min(t + 2, tlast), hv[P]);
            cslot = cslot == NSLOT - 1 ? 0 : cslot + 1;
            hslot = hslot == NSLOT - 1 ? 0 : hslot + 1;
            __syncthreads();

            if (active) {
                double cen[NC];
                static_for<0, NC>([&]<int C>() {
                    if constexpr (Staging<Sys, R>::staged(C)) cen[C] = q[Staging<Sys, R>::slot(C)][H];
                    else cen[C] = cv[P][C];
                });
                double sigma = 0.0;
                if (T > 0) {
                    const int zg = z + a.off[2];
                    const int dzv = zg - min(max(zg, T), a.dom[2] - 1 - T);
                    const int d2 = cxy2 + dzv * dzv;
                    sigma = d2 == 0 ? 0.0 : sig[d2];
                }
                const View<DOWN> v{q, reinterpret_cast<const Plane *>(buf), cen, a, lane + H, row + H};
                static_for<0, NC>([&]<int C>() {
                    double k = Sys::template user<N, C>(v, a);
                    if constexpr (R > 0) {
                        double ko_term = 0.0;
                        ko_term += KO<R, 0, C>(v);
                        ko_term += KO<R, 1, C>(v);
                        ko_term += KO<R, 2, C>(v);
                        k = k + a.ko_strength * ko_term;
                    }
                    if (T > 0) k = k - sigma * (cen[C] - a.target[C]);
                    const double initc = read_init ? iv[P][C] : cen[C];
                    double accv;
                    if constexpr (is_first) accv = initc + a.b * k;
                    else accv = av[P][C] + a.b * k;
                    mem.template put_acc<C>(accv, ch, zp);
                    if constexpr (write_argn) mem.template put_argn<C>(initc + a.a_next * k, ch, zp);
                });
            }
        };
        int t = 0;
        for (; t + 1 < cnt; t += 2) {
            step.template operator()<0>(t);
            step.template operator()<1>(t + 1);
        }
        if (t < cnt) step.template operator()<0>(t);
    }

    // Copy the ABL rate table (3 T^2 + 1 entries) into LDS; visible to the
    // block after the first barrier of run_tile.
    __device__ __forceinline__ static const double *stage_sigma(const double *tab, int T, double *dst) {
        if (T <= 0) return nullptr;
        const int nt = 3 * T * T + 1;
        for (int i = threadIdx.x; i < nt; i += blockDim.x) dst[i] = tab[i];
        return dst;
    }

    // single box: block -> (tile x, tile y, z chunk), optionally giving each
    // XCD (blocks are dealt to the 8 XCDs round robin) a contiguous run of tiles
    __device__ __forceinline__ static void run(const StageArgs &a) {
        __shared__ double lds[LDS_TOTAL];
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
        const double *sig = stage_sigma(a.sigma_tab, a.abl_T, sig_lds);
        // zalt 1: alternate the march direction per chunk so that the planes two
        // neighbouring chunks share are read by both at about the same time;
        // zalt 2: also reverse every direction in RK4 stages 1 and 3, so that a
        // stage starts at the planes the previous stage wrote last (stage 1
        // 1 % faster on wave256)
        const bool flip = a.zalt == 2 && (a.stage == kStage1 || a.stage == kStage3);
        if ((a.zalt != 0 && (zc & 1)) != flip) run_tile<true>(a, tix * TX, tiy * TY, z0, z1, lds, sig);
        else run_tile<false>(a, tix * TX, tiy * TY, z0, z1, lds, sig);
    }
};


template <class Sys, int N, int R, int TY, bool ADJ, int RS = 0, bool BUF = true, int TXW = 64, int SM = -1>
__global__ __launch_bounds__(TXW * TY) void fused_stage_kernel(const StageArgs a) {
    FusedStage<Sys, N, R, TY, ADJ, RS, BUF, TXW, SM>::run(a);
}

// Halo lookahead of the ring variant by tile height: 3 planes for 64x16 tiles
// (one block per CU; five plane buffers), 2 for 64x8 tiles (four buffers, so
// that two blocks share a CU's LDS).
template <int TY, int H> constexpr int ring_slots() { return TY >= 16 ? 3 : TY == 8 ? 2 : 0; }

// Tile heights with a halo-ring variant: use it when asked for and when its
// LDS (static + the dynamic ABL table) fits a block on this device.
template <class Sys, int N, int R, int TY, bool ADJ, int TXW = 64, int SM = -1>
hipError_t launch_tile(const StageArgs &a, dim3 grid, size_t shm, hipStream_t stream) {
    constexpr int RS = ring_slots<TY, FusedStage<Sys, N, R, TY, ADJ>::H>();
    using Ring = FusedStage<Sys, N, R, TY, ADJ, RS, true, TXW, SM>;
    if constexpr (RS > 0 && Ring::FITS) {
        if (a.halo_ring && sizeof(double) * Ring::LDS_TOTAL + shm <= max_block_lds()) {
            hipLaunchKernelGGL((fused_stage_kernel<Sys, N, R, TY, ADJ, RS, true, TXW, SM>), grid, dim3(TXW * TY),
                               shm, stream, a);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((fused_stage_kernel<Sys, N, R, TY, ADJ, 0, true, TXW, SM>), grid, dim3(TXW * TY), shm, stream,
                       a);
    return hipGetLastError();
}

// SM >= 0: stage-specific kernels for the 64x8, 64x16 and 32x32 tiles; for
// the other tile heights and the pointer-access kernel it returns
// hipErrorNotSupported without launching, and the caller uses the generic
// launcher (SM = -1).
template <class Sys, int N, int R, bool ADJ = false, int SM = -1>
hipError_t launch_fused(const StageArgs &a0, int tile_y, hipStream_t stream) {
    StageArgs a = a0;
    // registers of 4 GiB or more run one pointer-access kernel (64 x 8 tiles)
    const bool buf = !a.ptr_access && regs_fit_buffer(a, Sys::NC);
    if constexpr (SM >= 0)
        if (!buf || (tile_y != 8 && tile_y != 16 && tile_y != 32)) return hipErrorNotSupported;
    if (!buf) tile_y = 8;
    const int txw = tile_y == 32 ? 32 : 64; // tile_y 32: 32 x 32 tiles, two rows per wave
    a.tiles_x = (a.nx + txw - 1) / txw;
    a.tiles_y = (a.ny + tile_y - 1) / tile_y;
    a.nzc = (a.nz + a.zchunk - 1) / a.zchunk;
    const long long nblocks = (long long)a.tiles_x * a.tiles_y * a.nzc;
    if (nblocks <= 0 || nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)nblocks);
    const size_t shm = sigma_lds_bytes(a.abl_T);
    if constexpr (SM < 0) {
        if (!buf) {
            hipLaunchKernelGGL((fused_stage_kernel<Sys, N, R, 8, ADJ, 0, false>), grid, dim3(512), shm, stream, a);
            return hipGetLastError();
        }
        if (tile_y == 2) {
            hipLaunchKernelGGL((fused_stage_kernel<Sys, N, R, 2, ADJ>), grid, dim3(128), shm, stream, a);
            return hipGetLastError();
        }
        if (tile_y == 4) {
            hipLaunchKernelGGL((fused_stage_kernel<Sys, N, R, 4, ADJ>), grid, dim3(256), shm, stream, a);
            return hipGetLastError();
        }
    }
    switch (tile_y) {
    case 8: return launch_tile<Sys, N, R, 8, ADJ, 64, SM>(a, grid, shm, stream);
    case 16: return launch_tile<Sys, N, R, 16, ADJ, 64, SM>(a, grid, shm, stream);
    case 32: return launch_tile<Sys, N, R, 32, ADJ, 32, SM>(a, grid, shm, stream);
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Supported (stencil width, KO order) pairs: r in {0, n, n+1} with r <= 4.
template <class Sys, bool ADJ = false, int SM = -1> LaunchFn lookup_fused(int n, int r) {
#define TSP_CASE(NN, RR)                                                                                              \
    if (n == NN && r == RR) return &launch_fused<Sys, NN, RR, ADJ, SM>;
    TSP_CASE(1, 0)
    TSP_CASE(1, 1)
    TSP_CASE(1, 2)
    TSP_CASE(2, 0)
    TSP_CASE(2, 2)
    TSP_CASE(2, 3)
    TSP_CASE(3, 0)
    TSP_CASE(3, 3)
    TSP_CASE(3, 4)
    TSP_CASE(4, 0)
    TSP_CASE(4, 4)
#undef TSP_CASE
    return nullptr;
}

} // namespace tsp
