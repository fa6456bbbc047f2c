// level.hip -- device kernels of a level split into several boxes on one GPU
// (tsp_level_*): the ghost exchange between the boxes and the scatter/gather
// of the domain's valid state.
//
// The exchange is SingleLevelIntegrator::FillPatch (SingleLevelIntegrator.hpp:153-163)
// on a multi-box level, i.e. amrex::FillPatchSingleLevel's FillBoundary and
// physical-boundary steps: every ghost cell of every box whose domain index
// (after the periodic wrap along periodic axes) lies inside the domain takes
// the value of the valid cell of the box that owns that index (faces, edges
// and corners, neighbouring boxes and periodic images alike); a ghost cell
// outside the domain along a non-periodic axis takes the Dirichlet values
// (DirichletBC) or keeps its memory (NoOpBC). One thread per ghost cell, boxes
// along the grid's y dimension; the owner box is found through per-axis
// tables of the boxes' index ranges.
#include "batch_kernel.hpp"

namespace tsp {

__device__ __forceinline__ int owner_box(const LevelArgs &L, int x, int y, int z) {
    const int bx = L.axis_box[x], by = L.axis_box[L.dom[0] + y], bz = L.axis_box[L.dom[0] + L.dom[1] + z];
    return bx + L.nbox[0] * (by + L.nbox[1] * bz);
}

// ghost cell t of a box with valid extent (nx, ny, nz) and ghost width G: z-slabs,
// then y-slabs inside the valid z range, then x-slabs inside the valid y, z range
__device__ __forceinline__ void shell_cell(long long t, int nx, int ny, int nz, int G, int &i, int &j, int &k) {
    const int gx = nx + 2 * G, gy = ny + 2 * G;
    const long long na = 2LL * G * gy * gx;
    const long long nb = (long long)nz * 2 * G * gx;
    if (t < na) {
        const long long plane = (long long)gy * gx;
        const int kk = (int)(t / plane);
        const long long r = t % plane;
        j = (int)(r / gx) - G;
        i = (int)(r % gx) - G;
        k = kk < G ? kk - G : nz + (kk - G);
    } else if (t < na + nb) {
        const long long u = t - na;
        const long long plane = 2LL * G * gx;
        k = (int)(u / plane);
        const long long r = u % plane;
        const int jj = (int)(r / gx);
        i = (int)(r % gx) - G;
        j = jj < G ? jj - G : ny + (jj - G);
    } else {
        const long long u = t - na - nb;
        const long long plane = (long long)ny * 2 * G;
        k = (int)(u / plane);
        const long long r = u % plane;
        j = (int)(r / (2 * G));
        const int ii = (int)(r % (2 * G));
        i = ii < G ? ii - G : nx + (ii - G);
    }
}

__global__ __launch_bounds__(256) void level_exchange_kernel(const LevelArgs L, int role) {
    const int b = blockIdx.y;
    const BatchInstance &I = L.inst[b];
    const int G = L.G;
    const long long total =
        (long long)(I.nx + 2 * G) * (I.ny + 2 * G) * (I.nz + 2 * G) - (long long)I.nx * I.ny * I.nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i, j, k;
    shell_cell(t, I.nx, I.ny, I.nz, G, i, j, k);
    int g[3] = {I.off[0] + i, I.off[1] + j, I.off[2] + k};
    bool phys = false;
    for (int a = 0; a < 3; ++a)
        if (g[a] < 0 || g[a] >= L.dom[a]) {
            if (L.periodic[a]) g[a] += g[a] < 0 ? L.dom[a] : -L.dom[a];
            else phys = true;
        }
    double *dst = I.reg[role] + i + (long long)j * I.sj + (long long)k * I.sk;
    if (phys) {
        if (L.dirichlet)
            for (int c = 0; c < L.nc; ++c) dst[(long long)c * I.sc] = L.target[c];
        return;
    }
    const BatchInstance &S = L.inst[owner_box(L, g[0], g[1], g[2])];
    const double *src = S.reg[role] + (g[0] - S.off[0]) + (long long)(g[1] - S.off[1]) * S.sj +
                        (long long)(g[2] - S.off[2]) * S.sk;
    for (int c = 0; c < L.nc; ++c) dst[(long long)c * I.sc] = src[(long long)c * S.sc];
}

// the domain's dense valid state a[((c*NZ+z)*NY+y)*NX+x] into (scatter) or out of (gather) the boxes
template <bool SCATTER>
__global__ __launch_bounds__(256) void level_copy_kernel(const LevelArgs L, int role, double *dense) {
    const long long ncell = (long long)L.dom[0] * L.dom[1] * L.dom[2];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const int x = (int)(t % L.dom[0]);
    const int y = (int)((t / L.dom[0]) % L.dom[1]);
    const int z = (int)(t / ((long long)L.dom[0] * L.dom[1]));
    const BatchInstance &S = L.inst[owner_box(L, x, y, z)];
    double *reg = S.reg[role] + (x - S.off[0]) + (long long)(y - S.off[1]) * S.sj + (long long)(z - S.off[2]) * S.sk;
    for (int c = 0; c < L.nc; ++c) {
        if constexpr (SCATTER) reg[(long long)c * S.sc] = dense[c * ncell + t];
        else dense[c * ncell + t] = reg[(long long)c * S.sc];
    }
}

hipError_t launch_level_exchange(const LevelArgs &L, int role, long long max_shell, hipStream_t stream) {
    if (max_shell <= 0) return hipSuccess;
    const dim3 grid((unsigned)((max_shell + 255) / 256), (unsigned)L.nboxes);
    hipLaunchKernelGGL(level_exchange_kernel, grid, dim3(256), 0, stream, L, role);
    return hipGetLastError();
}

hipError_t launch_level_copy(const LevelArgs &L, int role, double *dense, bool scatter, hipStream_t stream) {
    const long long ncell = (long long)L.dom[0] * L.dom[1] * L.dom[2];
    const dim3 grid((unsigned)((ncell + 255) / 256));
    if (scatter) hipLaunchKernelGGL(level_copy_kernel<true>, grid, dim3(256), 0, stream, L, role, dense);
    else hipLaunchKernelGGL(level_copy_kernel<false>, grid, dim3(256), 0, stream, L, role, dense);
    return hipGetLastError();
}

} // namespace tsp
