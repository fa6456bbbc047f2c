// ghost.hip -- device ghost-cell fill and valid-region pack/unpack.
//
// The fill is SingleLevelIntegrator::FillPatch (SingleLevelIntegrator.hpp:153-163),
// i.e. amrex::FillPatchSingleLevel(mf, t, {&mf}, {t}, ...) on a single box:
// FillBoundary copies the periodic image of a valid cell into every ghost cell
// that lies outside the domain along periodic axes only; the physical-boundary
// functor then handles ghost cells outside the domain along a non-periodic
// axis: DirichletBC (bcs/DirichletBC.hpp:43-65) writes all boundary values,
// NoOpBC (bcs/NoOpBC.hpp:41-46) writes nothing. Faces, edges and corners are
// all covered. One thread per ghost cell; the shell is enumerated as z-slabs,
// then y-slabs inside the valid z range, then x-slabs inside the valid y,z range.
#include "common.hpp"

namespace tsp {

__global__ __launch_bounds__(256) void ghost_fill_kernel(const GhostArgs g, long long total) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int G = g.G;
    const int gx = g.nx + 2 * G, gy = g.ny + 2 * G;
    const long long na = 2LL * G * gy * gx;
    const long long nb = (long long)g.nz * 2 * G * gx;
    int i, j, k;
    if (t < na) {
        const long long plane = (long long)gy * gx;
        const int kk = (int)(t / plane);
        const long long r = t % plane;
        j = (int)(r / gx) - G;
        i = (int)(r % gx) - G;
        k = kk < G ? kk - G : g.nz + (kk - G);
    } else if (t < na + nb) {
        const long long u = t - na;
        const long long plane = 2LL * G * gx;
        k = (int)(u / plane);
        const long long r = u % plane;
        const int jj = (int)(r / gx);
        i = (int)(r % gx) - G;
        j = jj < G ? jj - G : g.ny + (jj - G);
    } else {
        const long long u = t - na - nb;
        const long long plane = (long long)g.ny * 2 * G;
        k = (int)(u / plane);
        const long long r = u % plane;
        j = (int)(r / (2 * G));
        const int ii = (int)(r % (2 * G));
        i = ii < G ? ii - G : g.nx + (ii - G);
    }
    int si = i, sj = j, sk = k;
    bool phys = false;
    if (si < 0 || si >= g.nx) {
        if (g.periodic[0]) si += si < 0 ? g.nx : -g.nx;
        else phys = true;
    }
    if (sj < 0 || sj >= g.ny) {
        if (g.periodic[1]) sj += sj < 0 ? g.ny : -g.ny;
        else phys = true;
    }
    if (sk < 0 || sk >= g.nz) {
        if (g.periodic[2]) sk += sk < 0 ? g.nz : -g.nz;
        else phys = true;
    }
    const long long dst = i + (long long)j * g.sj + (long long)k * g.sk;
    if (phys) {
        if (g.dirichlet)
            for (int c = 0; c < g.nc; ++c) g.p[dst + (long long)c * g.sc] = g.target[c];
    } else {
        const long long src = si + (long long)sj * g.sj + (long long)sk * g.sk;
        for (int c = 0; c < g.nc; ++c) g.p[dst + (long long)c * g.sc] = g.p[src + (long long)c * g.sc];
    }
}

hipError_t launch_ghost_fill(const GhostArgs &g, hipStream_t stream) {
    const int G = g.G;
    if (G == 0) return hipSuccess;
    const long long gx = g.nx + 2 * G, gy = g.ny + 2 * G;
    const long long total = 2LL * G * gy * gx + (long long)g.nz * 2 * G * gx + (long long)g.nz * g.ny * 2 * G;
    const long long blocks = (total + 255) / 256;
    hipLaunchKernelGGL(ghost_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, g, total);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void unpack_kernel(double *reg0, const double *dense, int nx, int ny, int nz,
                                                     int nc, int sj, int sk, int sc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nx * ny * nz * nc;
    if (t >= total) return;
    const int i = (int)(t % nx);
    long long r = t / nx;
    const int j = (int)(r % ny);
    r /= ny;
    const int k = (int)(r % nz);
    const int c = (int)(r / nz);
    reg0[i + (long long)j * sj + (long long)k * sk + (long long)c * sc] = dense[t];
}

__global__ __launch_bounds__(256) void pack_kernel(double *dense, const double *reg0, int nx, int ny, int nz, int nc,
                                                   int sj, int sk, int sc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nx * ny * nz * nc;
    if (t >= total) return;
    const int i = (int)(t % nx);
    long long r = t / nx;
    const int j = (int)(r % ny);
    r /= ny;
    const int k = (int)(r % nz);
    const int c = (int)(r / nz);
    dense[t] = reg0[i + (long long)j * sj + (long long)k * sk + (long long)c * sc];
}

hipError_t launch_unpack(double *reg0, const double *dense, int nx, int ny, int nz, int nc, int sj, int sk, int sc,
                         hipStream_t stream) {
    const long long total = (long long)nx * ny * nz * nc;
    hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, reg0, dense, nx, ny,
                       nz, nc, sj, sk, sc);
    return hipGetLastError();
}

hipError_t launch_pack(double *dense, const double *reg0, int nx, int ny, int nz, int nc, int sj, int sk, int sc,
                       hipStream_t stream) {
    const long long total = (long long)nx * ny * nz * nc;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, dense, reg0, nx, ny,
                       nz, nc, sj, sk, sc);
    return hipGetLastError();
}

} // namespace tsp

namespace tsp {

// Receiver trace of the adjoint solver. `counter` holds the trace row; the
// kernels read and advance it themselves so that captured graphs can be
// replayed step after step. One workgroup; the row is read before it is
// advanced by thread 0 after the barrier.
__global__ __launch_bounds__(256) void gather_trace_kernel(const double *phi, const int *rec_off, int nrec,
                                                           double *trace, int nrows, int *counter) {
    const int row = *counter;
    if (row >= 0 && row < nrows)
        for (int r = threadIdx.x; r < nrec; r += blockDim.x) trace[(long long)row * nrec + r] = phi[rec_off[r]];
    __syncthreads();
    if (threadIdx.x == 0) *counter = row + 1;
}

// lambda[rec] += dtrace[row][r], receivers in order (several may share a cell),
// then the row moves back by one.
__global__ void inject_trace_kernel(double *lam, const int *rec_off, int nrec, const double *dtrace, int nrows,
                                    int *counter) {
    if (threadIdx.x != 0) return;
    const int row = *counter;
    if (row >= 0 && row < nrows)
        for (int r = 0; r < nrec; ++r) lam[rec_off[r]] += dtrace[(long long)row * nrec + r];
    *counter = row - 1;
}

__global__ void set_counter_kernel(int *counter, int v) { *counter = v; }

__global__ __launch_bounds__(256) void zero_kernel(double *p, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

hipError_t launch_gather_trace(const double *phi, const int *rec_off, int nrec, double *trace, int nrows,
                               int *counter, hipStream_t stream) {
    hipLaunchKernelGGL(gather_trace_kernel, dim3(1), dim3(256), 0, stream, phi, rec_off, nrec, trace, nrows, counter);
    return hipGetLastError();
}

hipError_t launch_inject_trace(double *lam, const int *rec_off, int nrec, const double *dtrace, int nrows,
                               int *counter, hipStream_t stream) {
    hipLaunchKernelGGL(inject_trace_kernel, dim3(1), dim3(64), 0, stream, lam, rec_off, nrec, dtrace, nrows, counter);
    return hipGetLastError();
}

hipError_t launch_set_counter(int *counter, int v, hipStream_t stream) {
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(1), 0, stream, counter, v);
    return hipGetLastError();
}

hipError_t launch_zero(double *p, long long n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

} // namespace tsp
