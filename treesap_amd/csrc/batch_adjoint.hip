// batch_adjoint.hip -- batched kernels of the discrete adjoint: the reverse RK4
// stages of the transposed (phi, Pi) system for every variant group, and the
// receiver-trace gather/inject of all instances in one launch each. The
// reverse stage body is the single-box adjoint's tile body (FusedStage with
// ADJ), so every instance's gradient is bit-identical to its single-box one.
#include "batch_kernel.hpp"

namespace tsp {

hipError_t launch_batched_phipi_adjoint(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T,
                                        int halo_ring, hipStream_t stream) {
    hipError_t e = hipErrorInvalidValue;
    static_for<0, kBatchVariants>([&]<int V>() {
        if (variant == V)
            e = launch_batched<SysPhiPiT, kBatchN[V], kBatchR[V], -1, true>(b, tile_y, grid, max_abl_T, halo_ring,
                                                                           stream);
    });
    return e;
}

__global__ __launch_bounds__(256) void batch_gather_trace_kernel(const BatchInstance *inst, int role,
                                                                 const int *rec_off, int nrec, double *trace,
                                                                 int nrows, int *counter) {
    const int i = blockIdx.x;
    const int row = counter[i];
    const double *phi = inst[i].reg[role];
    if (row >= 0 && row < nrows)
        for (int r = threadIdx.x; r < nrec; r += blockDim.x)
            trace[((long long)i * nrows + row) * nrec + r] = phi[rec_off[(long long)i * nrec + r]];
    __syncthreads();
    if (threadIdx.x == 0) counter[i] = row + 1;
}

// lambda[rec] += dtrace[row][r], receivers in order (several may share a cell)
__global__ void batch_inject_trace_kernel(const BatchInstance *inst, int role, const int *rec_off, int nrec,
                                          const double *dtrace, int nrows, int *counter) {
    if (threadIdx.x != 0) return;
    const int i = blockIdx.x;
    const int row = counter[i];
    double *lam = inst[i].reg[role];
    if (row >= 0 && row < nrows)
        for (int r = 0; r < nrec; ++r)
            lam[rec_off[(long long)i * nrec + r]] += dtrace[((long long)i * nrows + row) * nrec + r];
    counter[i] = row - 1;
}

__global__ void batch_set_counter_kernel(int *counter, int count, int v) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) counter[i] = v;
}

hipError_t launch_batch_gather_trace(const BatchInstance *inst, int count, int role, const int *rec_off, int nrec,
                                     double *trace, int nrows, int *counter, hipStream_t stream) {
    hipLaunchKernelGGL(batch_gather_trace_kernel, dim3(count), dim3(256), 0, stream, inst, role, rec_off, nrec, trace,
                       nrows, counter);
    return hipGetLastError();
}

hipError_t launch_batch_inject_trace(const BatchInstance *inst, int count, int role, const int *rec_off, int nrec,
                                     const double *dtrace, int nrows, int *counter, hipStream_t stream) {
    hipLaunchKernelGGL(batch_inject_trace_kernel, dim3(count), dim3(64), 0, stream, inst, role, rec_off, nrec, dtrace,
                       nrows, counter);
    return hipGetLastError();
}

hipError_t launch_batch_set_counter(int *counter, int count, int v, hipStream_t stream) {
    hipLaunchKernelGGL(batch_set_counter_kernel, dim3(1), dim3(256), 0, stream, counter, count, v);
    return hipGetLastError();
}

} // namespace tsp
