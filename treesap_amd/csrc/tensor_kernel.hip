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
// six h components in increasing order can accumulate every sum in exactly
// the reference's order: each accumulator receives at most one term per
// component. The kernel evaluates each needed derivative of a component once
// and adds it to the accumulators that use it, which keeps the results
// bit-identical to evaluating the expressions term by term.
//
// One thread per cell (64 x 4 x 1 blocks, x fastest), reading the source
// register with filled ghosts (faces, edges and corners); the components are
// visited in a loop that is not unrolled, so each component's stencils are
// scheduled on their own.
#include "common.hpp"

namespace tsp {

namespace {

struct Cell {
    const double *p; // source at (i, j, k, component 0)
    int sj, sk, sc;
    const StageArgs &a;
    template <int AX> __device__ __forceinline__ int step() const { return AX == 0 ? 1 : AX == 1 ? sj : sk; }
};

// util/stencil_impl.hpp:206-222 at offset `o` from the cell
template <int N, int AX, int C>
__device__ __forceinline__ double d1(const Cell &x, int o) {
    double st = 0.0;
    static_for<1, N + 1>([&]<int S>() {
        constexpr double w = cx::w1(N, S);
        st += w * (x.p[o + C * x.sc + S * x.template step<AX>()] - x.p[o + C * x.sc - S * x.template step<AX>()]);
    });
    return x.a.gf1[AX] * st;
}

// util/stencil_impl.hpp:254-285: outer OUT of inner IN
template <int N, int OUT, int IN, int C> __device__ __forceinline__ double d2(const Cell &x) {
    if constexpr (OUT == IN) {
        constexpr double w0 = cx::w2(N, 0);
        double st = w0 * x.p[C * x.sc];
        static_for<1, N + 1>([&]<int S>() {
            constexpr double w = cx::w2(N, S);
            st += w * (x.p[C * x.sc + S * x.template step<OUT>()] + x.p[C * x.sc - S * x.template step<OUT>()]);
        });
        return x.a.gf2[OUT] * st;
    } else {
        double st = 0.0;
        static_for<1, N + 1>([&]<int S>() {
            constexpr double w = cx::w1(N, S);
            st += w * (d1<N, IN, C>(x, S * x.template step<OUT>()) - d1<N, IN, C>(x, -S * x.template step<OUT>()));
        });
        return x.a.gf1[OUT] * st;
    }
}

// util/stencil.hpp:77-104
template <int R, int AX, int C> __device__ __forceinline__ double ko(const Cell &x) {
    constexpr double c0 = static_cast<double>(cx::ko_coef(R, 0));
    constexpr double norm = cx::ko_norm(R);
    double result = c0 * x.p[C * x.sc];
    static_for<1, R + 1>([&]<int S>() {
        constexpr double cs = static_cast<double>(cx::ko_coef(R, S));
        result += cs * (x.p[C * x.sc + S * x.template step<AX>()] + x.p[C * x.sc - S * x.template step<AX>()]);
    });
    return result * norm * x.a.gf1[AX];
}

} // namespace

// component-indexed views of the accumulator tables (runtime c in 0..5)
__device__ __forceinline__ double pick3(int m, double v0, double v1, double v2) {
    return m == 0 ? v0 : m == 1 ? v1 : v2;
}

template <int N, int R>
__global__ __launch_bounds__(256) void tensor_stage_kernel(const StageArgs a) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int off = i + j * a.sj + k * a.sk;

    double sigma = 0.0;
    const int T = a.abl_T;
    if (T > 0) {
        const int dx = i - min(max(i, T), a.nx - 1 - T);
        const int dy = j - min(max(j, T), a.ny - 1 - T);
        const int dz = k - min(max(k, T), a.nz - 1 - T);
        const int d2v = dx * dx + dy * dy + dz * dz;
        sigma = d2v == 0 ? 0.0 : a.sigma_tab[d2v];
    }
    const int stage = a.stage;
    const bool is_first = stage == kStage0 || stage == kEuler;
    const bool write_argn = stage < kStage3;
    const bool read_init = write_argn && !is_first;

    // stage update of component c with right-hand side kv (SingleLevelRK4Integrator.hpp:142-154)
    auto update = [&](int c, double kv, double cenc) {
        const int o = off + c * a.sc;
        const double initc = read_init ? a.init[o] : cenc;
        if (is_first) a.acc[o] = initc + a.b * kv;
        else a.acc[o] = a.acc[o] + a.b * kv;
        if (write_argn) a.argn[o] = initc + a.a_next * kv;
    };
    auto decorate = [&](int c, double kv, const Cell &x, double cenc) {
        if constexpr (R > 0) {
            double t = 0.0;
            t += ko<R, 0, 0>(x);
            t += ko<R, 1, 0>(x);
            t += ko<R, 2, 0>(x);
            kv = kv + a.ko_strength * t;
        }
        if (T > 0) kv = kv - sigma * (cenc - a.target[c]);
        return kv;
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
#pragma unroll 1
    for (int c = 0; c < 6; ++c) {
        const Cell x{a.src + off + c * a.sc, a.sj, a.sk, a.sc, a};
        double dv[3][3];
        static_for<0, 3>([&]<int O>() { static_for<0, 3>([&]<int I>() { dv[O][I] = d2<N, O, I, 0>(x); }); });
        const int row = c >= 3 ? 2 : c >= 1 ? 1 : 0;
        const int col = c - (row * (row + 1)) / 2;
        // f[A][B] gets D_m D_A h_{mB} where sym2(m, B) == c, i.e. (m, B) is (row, col) or (col, row)
        static_for<0, 3>([&]<int A>() {
            static_for<0, 3>([&]<int B>() {
                if (B == col) f[A][B] += pick3(row, dv[0][A], dv[1][A], dv[2][A]);
                else if (B == row) f[A][B] += pick3(col, dv[0][A], dv[1][A], dv[2][A]);
                if (row == col && B <= A) tdd[A][B] += dv[A][B];
                if (B <= A && A == row && B == col) {
                    lap[A][B] += dv[0][0];
                    lap[A][B] += dv[1][1];
                    lap[A][B] += dv[2][2];
                }
            });
        });
        // d_t h_c = -2 K_c (+ KO, ABL)
        const double cenc = x.p[0];
        const double kv = decorate(c, -2.0 * x.p[6 * a.sc], x, cenc);
        update(c, kv, cenc);
    }
    // K components: symmetrise<a,b>(f) - 0.5 * (tdd + lap)
#pragma unroll 1
    for (int c = 6; c < 12; ++c) {
        const Cell x{a.src + off + c * a.sc, a.sj, a.sk, a.sc, a};
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
        const double cenc = x.p[0];
        const double kv = decorate(c, sym - 0.5 * (tv + lv), x, cenc);
        update(c, kv, cenc);
    }
}

template <int N, int R> hipError_t launch_tensor(const StageArgs &a, int tile_y, hipStream_t stream) {
    const int ty = 4;
    (void)tile_y;
    const dim3 grid((a.nx + 63) / 64, (a.ny + ty - 1) / ty, a.nz);
    hipLaunchKernelGGL((tensor_stage_kernel<N, R>), grid, dim3(64 * ty), 0, stream, a);
    return hipGetLastError();
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
