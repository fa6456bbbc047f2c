// init_expr.hip -- initial data from expression programs, evaluated on the
// device (tsp_init_expr, tsp_level_init_expr).
//
// The reference's FunctionInit (init/FunctionInit.hpp:34-101) compiles one
// amrex::Parser expression per component and SingleLevelIntegrator's
// initialise_impl (SingleLevelIntegrator.hpp:75-93) calls the executors on the
// device at every cell centre x = lo + (i + 0.5) dx. Here the caller (the
// drop-in FunctionInit) compiles each expression into a postfix program of
// tsp_expr_op, and one thread per valid cell runs every component's program
// on a small value stack. The host checks every program before launch: known
// codes, no stack underflow, at most TSP_EXPR_MAX_STACK values, exactly one
// value left.
#include "capi_internal.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace tsp {

struct InitArgs {
    const InitBox *boxes;
    const tsp_expr_op *ops;
    const int *comp_start; // nc + 1 entries
    int nc;
    int index_lo[3]; // the domain's first cell index (amrex::Box smallEnd)
    double lo[3], dx[3];
};

__device__ double run_program(const tsp_expr_op *ops, int begin, int end, double x, double y, double z) {
    double st[TSP_EXPR_MAX_STACK];
    int sp = 0;
    for (int p = begin; p < end; ++p) {
        const tsp_expr_op op = ops[p];
        switch (op.code) {
        case TSP_EXPR_NUM: st[sp++] = op.value; break;
        case TSP_EXPR_X: st[sp++] = x; break;
        case TSP_EXPR_Y: st[sp++] = y; break;
        case TSP_EXPR_Z: st[sp++] = z; break;
        case TSP_EXPR_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
        case TSP_EXPR_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
        case TSP_EXPR_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
        case TSP_EXPR_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
        case TSP_EXPR_POW: --sp; st[sp - 1] = pow(st[sp - 1], st[sp]); break;
        case TSP_EXPR_ATAN2: --sp; st[sp - 1] = atan2(st[sp - 1], st[sp]); break;
        case TSP_EXPR_MIN: --sp; st[sp - 1] = fmin(st[sp - 1], st[sp]); break;
        case TSP_EXPR_MAX: --sp; st[sp - 1] = fmax(st[sp - 1], st[sp]); break;
        case TSP_EXPR_NEG: st[sp - 1] = -st[sp - 1]; break;
        case TSP_EXPR_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
        case TSP_EXPR_EXP: st[sp - 1] = exp(st[sp - 1]); break;
        case TSP_EXPR_LOG: st[sp - 1] = log(st[sp - 1]); break;
        case TSP_EXPR_LOG10: st[sp - 1] = log10(st[sp - 1]); break;
        case TSP_EXPR_SIN: st[sp - 1] = sin(st[sp - 1]); break;
        case TSP_EXPR_COS: st[sp - 1] = cos(st[sp - 1]); break;
        case TSP_EXPR_TAN: st[sp - 1] = tan(st[sp - 1]); break;
        case TSP_EXPR_ASIN: st[sp - 1] = asin(st[sp - 1]); break;
        case TSP_EXPR_ACOS: st[sp - 1] = acos(st[sp - 1]); break;
        case TSP_EXPR_ATAN: st[sp - 1] = atan(st[sp - 1]); break;
        case TSP_EXPR_SINH: st[sp - 1] = sinh(st[sp - 1]); break;
        case TSP_EXPR_COSH: st[sp - 1] = cosh(st[sp - 1]); break;
        case TSP_EXPR_TANH: st[sp - 1] = tanh(st[sp - 1]); break;
        case TSP_EXPR_ABS: st[sp - 1] = fabs(st[sp - 1]); break;
        case TSP_EXPR_FLOOR: st[sp - 1] = floor(st[sp - 1]); break;
        case TSP_EXPR_CEIL: st[sp - 1] = ceil(st[sp - 1]); break;
        case TSP_EXPR_HEAVISIDE: st[sp - 1] = st[sp - 1] > 0.0 ? 1.0 : 0.0; break;
        default: break; // codes are checked on the host
        }
    }
    return st[0];
}

__global__ __launch_bounds__(256) void init_expr_kernel(const InitArgs a) {
    const InitBox B = a.boxes[blockIdx.y];
    const long long ncell = (long long)B.n[0] * B.n[1] * B.n[2];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const int i = (int)(t % B.n[0]);
    const int j = (int)((t / B.n[0]) % B.n[1]);
    const int k = (int)(t / ((long long)B.n[0] * B.n[1]));
    // SingleLevelIntegrator.hpp:85-87, i, j, k the cell's domain index
    const double x = a.lo[0] + (a.index_lo[0] + B.off[0] + i + 0.5) * a.dx[0];
    const double y = a.lo[1] + (a.index_lo[1] + B.off[1] + j + 0.5) * a.dx[1];
    const double z = a.lo[2] + (a.index_lo[2] + B.off[2] + k + 0.5) * a.dx[2];
    double *dst = B.origin + i + (long long)j * B.sj + (long long)k * B.sk;
    for (int c = 0; c < a.nc; ++c) dst[(long long)c * B.sc] = run_program(a.ops, a.comp_start[c], a.comp_start[c + 1], x, y, z);
}

} // namespace tsp

namespace tsp_host {

namespace {

// stack effect of each code: values popped, values pushed
bool stack_effect(int code, int &pop, int &push) {
    push = 1;
    if (code >= TSP_EXPR_NUM && code <= TSP_EXPR_Z) pop = 0;
    else if (code >= TSP_EXPR_ADD && code <= TSP_EXPR_MAX) pop = 2;
    else if (code >= TSP_EXPR_NEG && code <= TSP_EXPR_HEAVISIDE) pop = 1;
    else return false;
    return true;
}

} // namespace

tsp_status check_expr_programs(const tsp_expr_op *ops, const int32_t *comp_start, int nc) {
    if (!ops || !comp_start) return fail(TSP_INVALID_ARG, "ops or comp_start is NULL");
    if (comp_start[0] != 0) return fail(TSP_INVALID_ARG, "comp_start[0] must be 0");
    for (int c = 0; c < nc; ++c) {
        if (comp_start[c + 1] <= comp_start[c])
            return fail(TSP_INVALID_ARG, "component " + std::to_string(c) + " has an empty program");
        if (comp_start[c + 1] > TSP_EXPR_MAX_OPS) return fail(TSP_INVALID_ARG, "more than TSP_EXPR_MAX_OPS operations");
        int depth = 0;
        for (int p = comp_start[c]; p < comp_start[c + 1]; ++p) {
            int pop = 0, push = 0;
            if (!stack_effect(ops[p].code, pop, push))
                return fail(TSP_INVALID_ARG, "unknown code " + std::to_string(ops[p].code) + " in component " +
                                                 std::to_string(c));
            if (depth < pop) return fail(TSP_INVALID_ARG, "stack underflow in component " + std::to_string(c));
            depth += push - pop;
            if (depth > TSP_EXPR_MAX_STACK)
                return fail(TSP_UNSUPPORTED, "component " + std::to_string(c) + " needs more than " +
                                                 std::to_string(TSP_EXPR_MAX_STACK) + " stack values");
        }
        if (depth != 1) return fail(TSP_INVALID_ARG, "component " + std::to_string(c) + " leaves " +
                                                         std::to_string(depth) + " values, not 1");
    }
    return TSP_OK;
}

// Evaluate the programs into the boxes (device pointers in `boxes`) and wait.
tsp_status run_init_expr(const std::vector<tsp::InitBox> &boxes, const tsp_desc &domain, const tsp_expr_op *ops,
                         const int32_t *comp_start, int nc, const int32_t index_lo[3], hipStream_t stream) {
    tsp::InitArgs a{};
    a.nc = nc;
    for (int d = 0; d < 3; ++d) {
        a.index_lo[d] = index_lo ? index_lo[d] : 0;
        a.lo[d] = domain.prob_lo[d];
        a.dx[d] = (domain.prob_hi[d] - domain.prob_lo[d]) / (double)domain.n[d]; // Geometry::CellSize
    }
    const size_t nops = (size_t)comp_start[nc];
    const size_t bytes_ops = nops * sizeof(tsp_expr_op), bytes_cs = (size_t)(nc + 1) * sizeof(int),
                 bytes_box = boxes.size() * sizeof(tsp::InitBox);
    char *dev = nullptr;
    TSP_HIP(hipMalloc((void **)&dev, bytes_ops + bytes_cs + bytes_box + 64));
    tsp::InitBox *d_boxes = (tsp::InitBox *)dev;
    int *d_cs = (int *)(dev + bytes_box);
    tsp_expr_op *d_ops = (tsp_expr_op *)(dev + ((bytes_box + bytes_cs + 15) & ~size_t(15)));
    long long max_cells = 0;
    for (const auto &b : boxes) max_cells = std::max(max_cells, (long long)b.n[0] * b.n[1] * b.n[2]);
    hipError_t e = hipMemcpyAsync(d_boxes, boxes.data(), bytes_box, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_cs, comp_start, bytes_cs, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ops, ops, bytes_ops, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        a.boxes = d_boxes;
        a.ops = d_ops;
        a.comp_start = d_cs;
        const dim3 grid((unsigned)((max_cells + 255) / 256), (unsigned)boxes.size());
        hipLaunchKernelGGL(tsp::init_expr_kernel, grid, dim3(256), 0, stream, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(e, "init_expr");
    return TSP_OK;
}

} // namespace tsp_host
