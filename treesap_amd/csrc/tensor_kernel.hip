// tensor_kernel.hip -- RK4 stage / Euler kernel of the linearised ADM system
// (h_ij, K_ij): 12 components held by two symmetric rank-2 aliases
// (pde-model/tensor-alias/SymRank2Alias.hpp:25-53, components xx, yx, yy, zx,
// zy, zz; h at 0, K at 6), with
//   d_t h_ab = -2 K_ab
//   d_t K_ab = symmetrise<a,b>( contract_m D_m D_a h_mb )
//              - 0.5 * ( contract_m D_a D_b h_mm + contract_m D_m D_m h_ab )
// (symmetrise: util/tensor_operations.hpp:123-125; contract:
// util/tensor_operations.hpp:45-61; D_p D_q with p != q is the outer first
// derivative along p of the inner one along q, util/stencil_impl.hpp:269-285;
// p == q is the unmixed second derivative, :254-268), decorated by KO and ABL
// like every other system.
//
// Every contraction above is a left fold over m = 0, 1, 2, and the component
// it reads, sym2(m, b) or sym2(m, m), increases with m. So one pass over the
// six h components in increasing order accumulates every sum in exactly the
// reference's order: each accumulator receives at most one term per component.
//
// Tiling. A workgroup owns a 16 x 8 x 2 block of cells (one per thread; three
// such workgroups share a CU, so some compute while others wait at a barrier)
// and visits the h components one at a time. For each it
// stages the block plus an H-wide halo of the (ghost-filled) source in LDS.
// The three unmixed second derivatives and the KO term are read from there.
// The K components' KO terms are star stencils that need no block region:
// a separate z-marching kernel evaluates them just before (TensorKO below),
// and the block kernel reads them back at its K updates. For the six mixed ones,
// the inner first derivative g_i = D_i h is first evaluated at every point an
// outer stencil reads (a "slab": the block extended by n along the outer
// axis) and kept in LDS; the outer derivative then reads g_i. g_i at a shifted
// point is the same expression the reference evaluates there, so the result
// is bit-identical to the reference's nested stencil, at 2n instead of
// (2n)^2 reads per mixed derivative.
#include "common.hpp"

namespace tsp {

namespace {

// one of three values by a runtime index in 0..2, without an indexed array
__device__ __forceinline__ double pick3(int m, double v0, double v1, double v2) {
    return m == 0 ? v0 : m == 1 ? v1 : v2;
}

// Block shapes (x, y, z cells, one per thread): 16 x 8 x 2 and 8 x 8 x 4
// (256 threads, two workgroups per CU), 16 x 8 x 4 (512 threads, one).
template <int N, int R, int BX, int BY, int BZ> struct TensorTile {
    static constexpr int TX = BX, TY = BY, TZ = BZ, NT = TX * TY * TZ;
    static constexpr int H = N > R ? N : R;
    static constexpr int RX = TX + 2 * H, RY = TY + 2 * H, RZ = TZ + 2 * H, REG = RX * RY * RZ;

    // slab of g along axis GA extended by N along axis EA: dims and size
    template <int EA> static constexpr int ex = TX + (EA == 0 ? 2 * N : 0);
    template <int EA> static constexpr int ey = TY + (EA == 1 ? 2 * N : 0);
    template <int EA> static constexpr int ez = TZ + (EA == 2 ? 2 * N : 0);
    template <int EA> static constexpr int ssize = ex<EA> * ey<EA> * ez<EA>;

    // slabs: g_x along y (D_y D_x), g_y along x (D_x D_y), g_z along x (D_x D_z),
    // g_x along z (D_z D_x), g_y along z (D_z D_y), g_z along y (D_y D_z), each
    // with its own LDS, so one barrier separates all fills from all outer derivatives
    static constexpr int P1 = ssize<1> + 2 * ssize<0>;
    static constexpr int P2 = 2 * ssize<2> + ssize<1>;
    static constexpr int GBUF = P1 + P2;
    static constexpr int LDS_DOUBLES = REG + GBUF;

    static __device__ __forceinline__ int ridx(int rx, int ry, int rz) { return (rz * RY + ry) * RX + rx; }

    // g along axis GA at region offset o (util/stencil_impl.hpp:206-222)
    template <int GA> static __device__ __forceinline__ double g_at_off(const double *u, int o, const StageArgs &a) {
        constexpr int st = GA == 0 ? 1 : GA == 1 ? RX : RX * RY;
        double s = 0.0;
        static_for<1, N + 1>([&]<int S>() {
            constexpr double w = cx::w1(N, S);
            s += w * (u[o + S * st] - u[o - S * st]);
        });
        return a.gf1[GA] * s;
    }

    // The thread's share of a slab of g along GA extended by N on both sides
    // along EA: the slab coordinate along EA runs over [0, T_EA + 2N); the
    // thread computes coordinates t_EA, t_EA + T_EA, ... below that bound, the
    // others equal to its own cell's. All offsets are compile-time constants
    // from the thread's own region and slab positions.
    template <int GA, int EA> static __device__ __forceinline__ void fill_slab(const double *u, double *g, int tx,
                                                                              int ty, int tz, int own,
                                                                              const StageArgs &a) {
        constexpr int TE = EA == 0 ? TX : EA == 1 ? TY : TZ;
        constexpr int rst = EA == 0 ? 1 : EA == 1 ? RX : RX * RY; // region stride along EA
        constexpr int gst = EA == 0 ? 1 : EA == 1 ? ex<EA> : ex<EA> * ey<EA>; // slab stride along EA
        const int te = EA == 0 ? tx : EA == 1 ? ty : tz;
        const int g0 = (tz * ey<EA> + ty) * ex<EA> + tx; // slab coordinate t_EA along EA
        const int r0 = own - N * rst;                     // the region point of slab coordinate t_EA
        static_for<0, (TE + 2 * N + TE - 1) / TE>([&]<int K>() {
            if (K == 0 || te + K * TE < TE + 2 * N) {
                const int r = r0 + K * TE * rst;
                g[g0 + K * TE * gst] = g_at_off<GA>(u, r, a);
            }
        });
    }

    // outer first derivative along EA of a slab extended along EA, at the thread's own cell
    template <int EA> static __device__ __forceinline__ double outer(const double *g, int tx, int ty, int tz,
                                                                    const StageArgs &a) {
        constexpr int st = EA == 0 ? 1 : EA == 1 ? ex<EA> : ex<EA> * ey<EA>;
        const int o = ((tz + (EA == 2 ? N : 0)) * ey<EA> + ty + (EA == 1 ? N : 0)) * ex<EA> + tx + (EA == 0 ? N : 0);
        double s = 0.0;
        static_for<1, N + 1>([&]<int S>() {
            constexpr double w = cx::w1(N, S);
            s += w * (g[o + S * st] - g[o - S * st]);
        });
        return a.gf1[EA] * s;
    }

    // unmixed second derivative along AX at the own cell (util/stencil_impl.hpp:254-268)
    template <int AX> static __device__ __forceinline__ double d2(const double *u, int o, const StageArgs &a) {
        constexpr int st = AX == 0 ? 1 : AX == 1 ? RX : RX * RY;
        constexpr double w0 = cx::w2(N, 0);
        double s = w0 * u[o];
        static_for<1, N + 1>([&]<int S>() {
            constexpr double w = cx::w2(N, S);
            s += w * (u[o + S * st] + u[o - S * st]);
        });
        return a.gf2[AX] * s;
    }

    // Kreiss-Oliger term along AX at the own cell (util/stencil.hpp:77-104)
    template <int AX> static __device__ __forceinline__ double ko(const double *u, int o, const StageArgs &a) {
        constexpr int st = AX == 0 ? 1 : AX == 1 ? RX : RX * RY;
        constexpr double c0 = static_cast<double>(cx::ko_coef(R, 0));
        constexpr double norm = cx::ko_norm(R);
        double result = c0 * u[o];
        static_for<1, R + 1>([&]<int S>() {
            constexpr double cs = static_cast<double>(cx::ko_coef(R, S));
            result += cs * (u[o + S * st] + u[o - S * st]);
        });
        return result * norm * a.gf1[AX];
    }

    static __device__ void run(const StageArgs &a, double *lds) {
        double *u = lds;
        double *g = lds + REG;
        const int tid = threadIdx.x;
        const int tx = tid % TX, ty = (tid / TX) % TY, tz = tid / (TX * TY);
        const int tiles_x = (a.nx + TX - 1) / TX;
        // blocks are dealt to the 8 XCDs round robin; give each XCD a contiguous
        // run of tiles so that neighbours share an L2 for their common halo
        const int nb = gridDim.x;
        int bid = blockIdx.x;
        {
            const int per = nb / 8, rem = nb % 8, xcd = bid % 8;
            bid = xcd * per + (xcd < rem ? xcd : rem) + bid / 8;
        }
        // Each XCD's blocks (a contiguous range of bid after the remap) cover one
        // group of GS consecutive (x, y) tiles (x fastest) over all z, swept in
        // slabs of ZG = a.zchunk z tiles: a slab's GS * ZG blocks, about what
        // the XCD's CUs hold at once, share their halos in the XCD's L2, and the
        // next slab rereads the previous one's last planes from there.
        const int ZG = max(a.zchunk, 1);
        const int tiles_y = (a.ny + TY - 1) / TY;
        const int tiles_z = (a.nz + TZ - 1) / TZ;
        const int nxy = tiles_x * tiles_y;
        const int GS = (nxy + 7) / 8;
        const int grp = bid / (GS * tiles_z);                 // group of (x, y) tiles
        const int r = bid - grp * GS * tiles_z;               // index inside the group
        const int gsz = min(GS, nxy - grp * GS);              // tiles in this group (the last may hold fewer)
        const int slab = r / (gsz * ZG);
        const int r2 = r - slab * gsz * ZG;
        const int zn = min(ZG, tiles_z - slab * ZG);          // z tiles in this slab (the last may hold fewer)
        const int txy = grp * GS + r2 / zn;
        const int z0 = (slab * ZG + r2 % zn) * TZ;
        const int x0 = (txy % tiles_x) * TX;
        const int y0 = (txy / tiles_x) * TY;
        const int i = x0 + tx, j = y0 + ty, k = z0 + tz;
        const bool active = i < a.nx && j < a.ny && k < a.nz;
        const int off = active ? i + j * a.sj + k * a.sk : 0;
        const int own = ridx(tx + H, ty + H, tz + H);

        const int stage = a.stage;
        const bool is_first = stage == kStage0 || stage == kEuler;
        const bool write_argn = stage < kStage3;
        const bool read_init = write_argn && !is_first;

        double sigma = 0.0;
        const int T = a.abl_T;
        if (T > 0 && active) {
            const int dx = i - min(max(i, T), a.nx - 1 - T);
            const int dy = j - min(max(j, T), a.ny - 1 - T);
            const int dz = k - min(max(k, T), a.nz - 1 - T);
            const int d2v = dx * dx + dy * dy + dz * dz;
            sigma = d2v == 0 ? 0.0 : a.sigma_tab[d2v];
        }

        // The source of one component over the block plus halo (the register's
        // ghosts are filled; indices of partial blocks are clamped into it) is
        // loaded into registers one component ahead and written to LDS at the
        // start of that component's pass, so its latency overlaps the previous
        // component's arithmetic.
        // Not every region point is read: no stencil reaches a corner of the
        // region (outside the block along all three axes). The point set is a
        // bit mask over the thread's points.
        constexpr int RPT = (REG + NT - 1) / NT;
        int roff[RPT];
        unsigned hmask = 0;
#pragma unroll
        for (int m = 0; m < RPT; ++m) {
            const int p = min(tid + m * NT, REG - 1);
            const int rx = p % RX, q = p / RX, ry = q % RY, rz = q / RY;
            const int x = min(x0 + rx - H, a.nx + a.G - 1);
            const int y = min(y0 + ry - H, a.ny + a.G - 1);
            const int z = min(z0 + rz - H, a.nz + a.G - 1);
            roff[m] = x + y * a.sj + z * a.sk;
            const int outside = (rx < H || rx >= H + TX) + (ry < H || ry >= H + TY) + (rz < H || rz >= H + TZ);
            if (tid + m * NT < REG && outside < 3) hmask |= 1u << m;
        }
        double pre[RPT];
        auto load_region = [&](int c, double (&dst)[RPT]) {
            const double *src = a.src + (long long)c * a.sc;
#pragma unroll
            for (int m = 0; m < RPT; ++m)
                if (hmask >> m & 1u) dst[m] = src[roff[m]];
        };
        auto stage_region = [&](const double (&from)[RPT], double *buf) {
#pragma unroll
            for (int m = 0; m < RPT; ++m)
                if (hmask >> m & 1u) buf[tid + m * NT] = from[m];
        };
        // stage update of component c with right-hand side kv (SingleLevelRK4Integrator.hpp:142-154);
        // the stores stream past the caches (non-temporal) that hold the regions' halos
        auto update = [&](int c, double kv, double cenc, double initv, double accv) {
            const long long o = off + (long long)c * a.sc;
            const double initc = read_init ? initv : cenc;
            if (is_first) __builtin_nontemporal_store(initc + a.b * kv, &a.acc[o]);
            else __builtin_nontemporal_store(accv + a.b * kv, &a.acc[o]);
            if (write_argn) __builtin_nontemporal_store(initc + a.a_next * kv, &a.argn[o]);
        };
        auto decorate = [&](int c, double kv, double ko_term, double cenc) {
            if constexpr (R > 0) kv = kv + a.ko_strength * ko_term;
            if (T > 0) kv = kv - sigma * (cenc - a.target[c]);
            return kv;
        };
        auto ko_sum = [&](const double *w) {
            double t = 0.0;
            if constexpr (R > 0) {
                t += ko<0>(w, own, a);
                t += ko<1>(w, own, a);
                t += ko<2>(w, own, a);
            }
            return t;
        };

        // f[A][B] = sum_m D_m D_A h_{mB}, tdd[A][B] = sum_m D_A D_B h_mm, lap[A][B] = sum_m D_m D_m h_AB
        double f[3][3], tdd[3][3], lap[3][3];
        static_for<0, 3>([&]<int A>() {
            static_for<0, 3>([&]<int B>() {
                f[A][B] = 0.0;
                tdd[A][B] = 0.0;
                lap[A][B] = 0.0;
            });
        });

        // h components in increasing order (the reference's contraction order)
        load_region(0, pre);
#pragma unroll 1
        for (int c = 0; c < 6; ++c) {
            __syncthreads(); // the previous component's reads of u and g are done
            stage_region(pre, u);
            // centre values first, then the prefetch: the memory counter
            // completes in order, so the update's wait for the centre values
            // then leaves the next region's loads in flight
            double kc = 0.0, initv = 0.0, accv = 0.0;
            if (active) {
                kc = a.src[off + (long long)(6 + c) * a.sc];
                if (read_init) initv = a.init[off + (long long)c * a.sc];
                if (!is_first) accv = a.acc[off + (long long)c * a.sc];
            }
            if (c < 5) load_region(c + 1, pre);
            __syncthreads();

            double dv[3][3];
            dv[0][0] = d2<0>(u, own, a);
            dv[1][1] = d2<1>(u, own, a);
            dv[2][2] = d2<2>(u, own, a);
            const double ko_term = ko_sum(u);
            const double cenc = u[own];

            const int row = c >= 3 ? 2 : c >= 1 ? 1 : 0;
            const int col = c - (row * (row + 1)) / 2;
            // Mixed derivatives this component feeds (outer p, inner q): f takes
            // D_row D_q and D_col D_q for every q, tdd takes D_p D_q (q < p) of
            // the diagonal components. The others are not evaluated:
            // D_y D_x (c != 3), D_x D_y and D_x D_z (c in {0, 1, 3}),
            // D_z D_x and D_z D_y (c != 1), D_y D_z (c in {1, 2, 4}).
            const bool need_yx = c != 3, need_xy = c <= 1 || c == 3, need_zx = c != 1;
            const bool need_yz = c == 1 || c == 2 || c == 4;

            // slabs of g_x along y, g_y along x, g_z along x, g_x along z, g_y along z, g_z along y
            constexpr int SA = ssize<1>, SB = ssize<0>, SD = ssize<2>;
            double *g2 = g + P1;
            if (need_yx) fill_slab<0, 1>(u, g, tx, ty, tz, own, a);
            if (need_xy) {
                fill_slab<1, 0>(u, g + SA, tx, ty, tz, own, a);
                fill_slab<2, 0>(u, g + SA + SB, tx, ty, tz, own, a);
            }
            if (need_zx) {
                fill_slab<0, 2>(u, g2, tx, ty, tz, own, a);
                fill_slab<1, 2>(u, g2 + SD, tx, ty, tz, own, a);
            }
            if (need_yz) fill_slab<2, 1>(u, g2 + 2 * SD, tx, ty, tz, own, a);
            __syncthreads();
            dv[1][0] = need_yx ? outer<1>(g, tx, ty, tz, a) : 0.0;
            dv[0][1] = need_xy ? outer<0>(g + SA, tx, ty, tz, a) : 0.0;
            dv[0][2] = need_xy ? outer<0>(g + SA + SB, tx, ty, tz, a) : 0.0;
            dv[2][0] = need_zx ? outer<2>(g2, tx, ty, tz, a) : 0.0;
            dv[2][1] = need_zx ? outer<2>(g2 + SD, tx, ty, tz, a) : 0.0;
            dv[1][2] = need_yz ? outer<1>(g2 + 2 * SD, tx, ty, tz, a) : 0.0;

            static_for<0, 3>([&]<int A>() {
                const double by_row = pick3(row, dv[0][A], dv[1][A], dv[2][A]);
                const double by_col = pick3(col, dv[0][A], dv[1][A], dv[2][A]);
                static_for<0, 3>([&]<int B>() {
                    if (B == col) f[A][B] += by_row;
                    else if (B == row) f[A][B] += by_col;
                    if (row == col && B <= A) tdd[A][B] += dv[A][B];
                    if (B <= A && A == row && B == col) {
                        lap[A][B] += dv[0][0];
                        lap[A][B] += dv[1][1];
                        lap[A][B] += dv[2][2];
                    }
                });
            });
            // d_t h_c = -2 K_c (+ KO, ABL)
            if (active) update(c, decorate(c, -2.0 * kc, ko_term, cenc), cenc, initv, accv);
        }

        // K components: symmetrise<a,b>(f) - 0.5 * (tdd + lap), plus the KO
        // sums TensorKO left in a.kko; their centre, init, acc and KO values
        // are loaded together (unrolled loop)
#pragma unroll
        for (int c = 6; c < 12; ++c) {
            double ko_term = 0.0, cenc = 0.0, initv = 0.0, accv = 0.0;
            if (active) {
                cenc = a.src[off + (long long)c * a.sc];
                if (read_init) initv = a.init[off + (long long)c * a.sc];
                if (!is_first) accv = a.acc[off + (long long)c * a.sc];
                if constexpr (R > 0) ko_term = a.kko[off + (long long)(c - 6) * a.sc];
            }
            const int row = c >= 9 ? 2 : c >= 7 ? 1 : 0;
            const int col = (c - 6) - (row * (row + 1)) / 2;
            double fab = 0.0, fba = 0.0, tv = 0.0, lv = 0.0;
            static_for<0, 3>([&]<int A>() {
                static_for<0, 3>([&]<int B>() {
                    if (A == row && B == col) {
                        fab = f[A][B];
                        tv = tdd[A][B];
                        lv = lap[A][B];
                    }
                    if (A == col && B == row) fba = f[A][B];
                });
            });
            const double sym = 0.5 * (fab + fba);
            if (active) update(c, decorate(c, sym - 0.5 * (tv + lv), ko_term, cenc), cenc, initv, accv);
        }
    }
};

// Kreiss-Oliger sums ((0 + KO_x) + KO_y) + KO_z of the six K components
// (KODecorator.hpp:58-70, stencil.hpp:77-104), written to a.kko (register
// layout) for the block kernel's K updates. These star stencils need no block
// region, so they stream: a workgroup of 64 x KY threads owns a 64 x KY
// column tile and marches a chunk of zc planes, KG components at a time. Each
// thread keeps its column's 2R + 1 planes of each component in a register
// queue (the z stencil); a plane's x and y neighbours pass through LDS, in two
// plane buffers used alternately, so one barrier per plane suffices. The next
// plane's halo values and queue heads are loaded while a plane computes
// (loading them two planes ahead, into two alternating register sets,
// measured 15 % slower per stage).
template <int R, int KY, int KG> struct TensorKO {
    static constexpr int LX = 64 + 2 * R, LY = KY + 2 * R, PL = LX * LY, NT = 64 * KY;
    static constexpr int LDS_DOUBLES = 2 * KG * PL;
    static_assert(R >= 1 && KY >= R, "the y halo rows are loaded by the first and last R rows of threads");
    static_assert(6 % KG == 0, "components are taken in equal groups");

    // the KO term along AX from the centre value c and the pair sums
    // u(+s) + u(-s), given by a functor of s (same operation order as TensorTile::ko)
    template <int AX, class F> static __device__ __forceinline__ double ko(double c, F &&at, const StageArgs &a) {
        constexpr double c0 = static_cast<double>(cx::ko_coef(R, 0));
        constexpr double norm = cx::ko_norm(R);
        double result = c0 * c;
        static_for<1, R + 1>([&]<int S>() {
            constexpr double cs = static_cast<double>(cx::ko_coef(R, S));
            result += cs * at(S);
        });
        return result * norm * a.gf1[AX];
    }

    static __device__ void run(const StageArgs &a, int zc, double *lds) {
        const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
        const int tiles_x = (a.nx + 63) / 64, tiles_y = (a.ny + KY - 1) / KY;
        // blocks are dealt to the 8 XCDs round robin; give each XCD a contiguous
        // run of tiles (x fastest, then y, then z chunk) so that neighbouring
        // tiles read their common y halo rows from one L2
        const int nb = gridDim.x;
        int bid = blockIdx.x;
        {
            const int per = nb / 8, rem = nb % 8, xcd = bid % 8;
            bid = xcd * per + (xcd < rem ? xcd : rem) + bid / 8;
        }
        const int txy = bid % (tiles_x * tiles_y), ch = bid / (tiles_x * tiles_y);
        const int i = (txy % tiles_x) * 64 + tx, j = (txy / tiles_x) * KY + ty;
        const int z0 = ch * zc, zend = min(z0 + zc, a.nz);
        const bool active = i < a.nx && j < a.ny;
        // Load positions are clamped into the ghost-filled allocation
        // ([-G, n + G) per axis): the lanes of a partial tile past n + G - 1
        // feed no active lane's stencil.
        const int xl = a.nx + a.G - 1, yl = a.ny + a.G - 1;
        const int ic = min(i, xl), jc = min(j, yl);
        const int cxy = ic + jc * a.sj; // the thread's column
        const bool hxlo = tx < R, hxhi = tx >= 64 - R, hylo = ty < R, hyhi = ty >= KY - R;
        const int hx = (hxlo ? i - R : min(i + R, xl)) + jc * a.sj; // x halo point (lanes within R of an x edge)
        const int hyl = ic + (j - R) * a.sj;                           // y halo points (rows within R of a y edge)
        const int hyh = ic + min(j + R, yl) * a.sj;
        // LDS slots of a plane buffer [LY][LX]: the own cell, then the halo points
        const int own = (ty + R) * LX + tx + R;
        const int sx = (ty + R) * LX + (hxlo ? tx : tx + 2 * R);
        const int syl = ty * LX + tx + R, syh = (ty + 2 * R) * LX + tx + R;
        const long long sk = a.sk;

#pragma unroll 1
        for (int g0 = 0; g0 < 6; g0 += KG) {
            if (g0 > 0) __syncthreads(); // the previous group's reads of the plane buffers are done
            const double *src[KG];
#pragma unroll
            for (int g = 0; g < KG; ++g) src[g] = a.src + (long long)(6 + g0 + g) * a.sc;
            double q[KG][2 * R + 1], vx[KG], vyl[KG], vyh[KG];
#pragma unroll
            for (int g = 0; g < KG; ++g) {
#pragma unroll
                for (int m = 0; m <= 2 * R; ++m) q[g][m] = src[g][cxy + (z0 - R + m) * sk];
                if (hxlo || hxhi) vx[g] = src[g][hx + z0 * sk];
                if (hylo) vyl[g] = src[g][hyl + z0 * sk];
                if (hyhi) vyh[g] = src[g][hyh + z0 * sk];
            }
#pragma unroll 1
            for (int z = z0; z < zend; ++z) {
                double *L = lds + ((z - z0) & 1) * (KG * PL);
#pragma unroll
                for (int g = 0; g < KG; ++g) {
                    double *Lg = L + g * PL;
                    Lg[own] = q[g][R];
                    if (hxlo || hxhi) Lg[sx] = vx[g];
                    if (hylo) Lg[syl] = vyl[g];
                    if (hyhi) Lg[syh] = vyh[g];
                }
                __syncthreads();
                const bool more = z + 1 < zend;
                if (more) {
#pragma unroll
                    for (int g = 0; g < KG; ++g) {
                        if (hxlo || hxhi) vx[g] = src[g][hx + (z + 1) * sk];
                        if (hylo) vyl[g] = src[g][hyl + (z + 1) * sk];
                        if (hyhi) vyh[g] = src[g][hyh + (z + 1) * sk];
                    }
                }
#pragma unroll
                for (int g = 0; g < KG; ++g) {
                    const double *Lg = L + g * PL;
                    const double c = q[g][R];
                    const double kx = ko<0>(c, [&](int s) { return Lg[own + s] + Lg[own - s]; }, a);
                    const double ky = ko<1>(c, [&](int s) { return Lg[own + s * LX] + Lg[own - s * LX]; }, a);
                    const double kz = ko<2>(c, [&](int s) { return q[g][R + s] + q[g][R - s]; }, a);
                    double sum = 0.0;
                    sum += kx;
                    sum += ky;
                    sum += kz;
                    if (active) __builtin_nontemporal_store(sum, &a.kko[i + (long long)j * a.sj + z * sk + (long long)(g0 + g) * a.sc]);
                }
#pragma unroll
                for (int g = 0; g < KG; ++g) {
#pragma unroll
                    for (int m = 0; m < 2 * R; ++m) q[g][m] = q[g][m + 1];
                    if (more) q[g][2 * R] = src[g][cxy + (z + 1 + R) * sk];
                }
            }
        }
    }
};

// z tiles of the KO kernel's column chunks, its tile height and component group
constexpr int kKoChunk = 32, kKoTileY = 8, kKoGroup = 2;

template <int R, int KY, int KG> __global__ __launch_bounds__(64 * KY) void tensor_ko_kernel(const StageArgs a, int zc) {
    __shared__ double lds[TensorKO<R, KY, KG>::LDS_DOUBLES];
    TensorKO<R, KY, KG>::run(a, zc, lds);
}

template <int R> hipError_t launch_tensor_ko(const StageArgs &a, hipStream_t stream) {
    using KO = TensorKO<R, kKoTileY, kKoGroup>;
    if (!a.kko || a.G < R) return hipErrorInvalidValue;
    const int zc = min(kKoChunk, a.nz);
    const long long blocks =
        (long long)((a.nx + 63) / 64) * ((a.ny + kKoTileY - 1) / kKoTileY) * ((a.nz + zc - 1) / zc);
    if (blocks <= 0 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((tensor_ko_kernel<R, kKoTileY, kKoGroup>), dim3((unsigned)blocks), dim3(KO::NT), 0, stream, a,
                       zc);
    return hipGetLastError();
}

// 256-thread blocks with stencils up to n = 3 are held to 168 VGPRs, three
// waves per SIMD (three workgroups per CU by their LDS): the compiler spills
// almost nothing there, and the extra wave per SIMD hides more of the barrier
// and load waits than the tighter allocation costs (measured 1.17e9 against
// 9.8e8 cell-updates/s on the 192^3 workload at the compiler's own 202
// VGPRs). At n = 4 the bound would spill 40 VGPRs.
template <int N, int R, int BX, int BY, int BZ>
__global__ __launch_bounds__((TensorTile<N, R, BX, BY, BZ>::NT))
__attribute__((amdgpu_waves_per_eu(TensorTile<N, R, BX, BY, BZ>::NT == 256 && N <= 3 ? 3 : 1))) void tensor_stage_kernel(
    const StageArgs a) {
    __shared__ double lds[TensorTile<N, R, BX, BY, BZ>::LDS_DOUBLES];
    TensorTile<N, R, BX, BY, BZ>::run(a, lds);
}

template <int N, int R, int BX, int BY, int BZ> hipError_t launch_tensor_shape(const StageArgs &a, hipStream_t stream) {
    using TT = TensorTile<N, R, BX, BY, BZ>;
    const long long blocks = (long long)((a.nx + TT::TX - 1) / TT::TX) * ((a.ny + TT::TY - 1) / TT::TY) *
                             ((a.nz + TT::TZ - 1) / TT::TZ);
    if (blocks <= 0 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if constexpr (R > 0) {
        const hipError_t e = launch_tensor_ko<R>(a, stream);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((tensor_stage_kernel<N, R, BX, BY, BZ>), dim3((unsigned)blocks), dim3(TT::NT), 0, stream, a);
    return hipGetLastError();
}

} // namespace

// The block shape by tile_y: 4 -> 8 x 8 x 4, 16 -> 16 x 8 x 4, otherwise 16 x 8 x 2.
template <int N, int R> hipError_t launch_tensor(const StageArgs &a, int tile_y, hipStream_t stream) {
    if (tile_y == 4) return launch_tensor_shape<N, R, 8, 8, 4>(a, stream);
    if (tile_y == 16) return launch_tensor_shape<N, R, 16, 8, 4>(a, stream);
    return launch_tensor_shape<N, R, 16, 8, 2>(a, stream);
}

LaunchFn lookup_tensor(int n, int r) {
#define TSP_TCASE(NN, RR)                                                                                             \
    if (n == NN && r == RR) return &launch_tensor<NN, RR>;
    TSP_TCASE(1, 0)
    TSP_TCASE(1, 1)
    TSP_TCASE(2, 0)
    TSP_TCASE(2, 2)
    TSP_TCASE(3, 0)
    TSP_TCASE(3, 3)
    TSP_TCASE(4, 0)
    TSP_TCASE(4, 4)
#undef TSP_TCASE
    return nullptr;
}

} // namespace tsp
