// adjoint_nl.hip -- the nonlinear part of the discrete adjoint of the
// NonLinear (phi, pi) system (examples/02-temporaries/NonLinearRHS.hpp:52-71).
//
// F_pi = Lap phi - gamma pi |g|^2 with g = D1(phi). The Jacobian at a stage
// argument (phi_b, pi_b) is the linear (phi, Pi) operator L plus N,
//   (N v)_pi = -gamma (v_pi |g|^2 + 2 pi_b g . D1(v_phi)),
// so every reverse stage adds N^T mu to nu = L^T mu (which the transposed
// stage kernel has already used twice: acc += nu, next mu = b lambda' + a nu):
//   (N^T mu)_pi  = -(gamma |g|^2) mu_pi,
//   (N^T mu)_phi = -sum_a gf1_a sum_s w1(n,s) (W_a(i - s e_a) - W_a(i + s e_a)),
//   W_a = ((2 gamma) pi_b) g_a mu_pi
// (W zero outside the box along non-periodic axes, periodic images along
// periodic ones). Two passes of one thread per valid cell: W and the pi part
// into a dense scratch, then the phi part gathered from W and both added to
// acc and next mu.
#include "common.hpp"

namespace tsp {

namespace {

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : v >= n ? v - n : v; }

template <int N> __global__ __launch_bounds__(256) void nl_adjoint_w_kernel(const NlAdjArgs a) {
    const long long cells = (long long)a.nx * a.ny * a.nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    const int i = (int)(t % a.nx), j = (int)((t / a.nx) % a.ny), k = (int)(t / ((long long)a.nx * a.ny));
    const int idx[3] = {i, j, k};
    const int nn[3] = {a.nx, a.ny, a.nz};
    double g[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        double st = 0.0;
        static_for<1, N + 1>([&]<int S>() {
            int p[3] = {i, j, k}, m[3] = {i, j, k};
            p[ax] = idx[ax] + S;
            m[ax] = idx[ax] - S;
            if (a.periodic[ax]) { // the forward stage reads periodic images there
                p[ax] = wrap(p[ax], nn[ax]);
                m[ax] = wrap(m[ax], nn[ax]);
            }
            const double fp = a.abar[p[0] + (long long)p[1] * a.sj + (long long)p[2] * a.sk];
            const double fm = a.abar[m[0] + (long long)m[1] * a.sj + (long long)m[2] * a.sk];
            st += cx::w1(N, S) * (fp - fm);
        });
        g[ax] = a.gf1[ax] * st;
    }
    double gsq = g[0] * g[0];
    gsq += g[1] * g[1];
    gsq += g[2] * g[2];
    const long long c0 = i + (long long)j * a.sj + (long long)k * a.sk;
    const double pib = a.abar[c0 + a.sc];
    const double mupi = a.mu_scale * a.mu[c0 + a.sc];
    const double tp = (2.0 * a.gamma) * pib;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) a.w[ax * cells + t] = (tp * g[ax]) * mupi;
    a.w[3 * cells + t] = -((a.gamma * gsq) * mupi);
}

template <int N> __global__ __launch_bounds__(256) void nl_adjoint_apply_kernel(const NlAdjArgs a) {
    const long long cells = (long long)a.nx * a.ny * a.nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    const int i = (int)(t % a.nx), j = (int)((t / a.nx) % a.ny), k = (int)(t / ((long long)a.nx * a.ny));
    const int idx[3] = {i, j, k};
    const int nn[3] = {a.nx, a.ny, a.nz};
    auto w_at = [&](int ax, int q[3]) -> double {
        for (int d = 0; d < 3; ++d)
            if (q[d] < 0 || q[d] >= nn[d]) {
                if (!a.periodic[d]) return 0.0;
                q[d] = wrap(q[d], nn[d]);
            }
        return a.w[ax * cells + q[0] + (long long)q[1] * a.nx + (long long)q[2] * a.nx * a.ny];
    };
    double cphi = 0.0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        double st = 0.0;
        static_for<1, N + 1>([&]<int S>() {
            int m[3] = {i, j, k}, p[3] = {i, j, k};
            m[ax] = idx[ax] - S;
            p[ax] = idx[ax] + S;
            st += cx::w1(N, S) * (w_at(ax, m) - w_at(ax, p));
        });
        cphi += a.gf1[ax] * st;
    }
    const double c_phi = -cphi;
    const double c_pi = a.w[3 * cells + t];
    const long long c0 = i + (long long)j * a.sj + (long long)k * a.sk;
    a.acc[c0] = a.acc[c0] + c_phi;
    a.acc[c0 + a.sc] = a.acc[c0 + a.sc] + c_pi;
    if (a.argn) {
        a.argn[c0] = a.argn[c0] + a.a_next * c_phi;
        a.argn[c0 + a.sc] = a.argn[c0 + a.sc] + a.a_next * c_pi;
    }
}

template <int N> hipError_t launch_n(const NlAdjArgs &a, hipStream_t stream) {
    const long long cells = (long long)a.nx * a.ny * a.nz;
    const dim3 grid((unsigned)((cells + 255) / 256));
    hipLaunchKernelGGL(nl_adjoint_w_kernel<N>, grid, dim3(256), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nl_adjoint_apply_kernel<N>, grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

} // namespace

hipError_t launch_nonlinear_adjoint_terms(const NlAdjArgs &a, hipStream_t stream) {
    switch (a.n) {
    case 1: return launch_n<1>(a, stream);
    case 2: return launch_n<2>(a, stream);
    case 3: return launch_n<3>(a, stream);
    case 4: return launch_n<4>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace tsp
