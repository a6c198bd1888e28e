// vgx_adapt.hip -- AdaptiveMap.adapt (src/vegas/_vegas.pyx:467-594) on device.
//
// Per axis: average the training data, smooth and compress it (alpha > 0),
// then invert the cumulative distribution onto new_ninc equal-weight
// increments.  Two implementations:
//   exact     one lane walks each axis with the reference's order of
//             operations (sequential sums, the while-loop CDF walk);
//   parallel  one 1024-lane workgroup per axis: tree reductions, an LDS
//             prefix sum and a binary search per new node.  Node i lands in
//             the first old increment j with prefix P_j >= i*F, exactly the
//             increment the sequential walk stops in (its running
//             accumulator equals P_j - i*F); only rounding differs.
#include "vgx_adapt.h"

namespace vgx {

// Fill the padding of one output row and nothing else.
__device__ inline void pad_row(double* ng, int nn, int new_ld) {
    for (int i = nn + 1; i < new_ld; ++i) ng[i] = ng[nn];
}

__device__ void adapt_axis_exact(const AdaptArgs& a, int d, double* avg, double* tmp, const double* g) {
    const int old = a.ninc[d], nn = a.new_ninc[d];
    double* ng = a.new_grid + (int64_t)d * a.new_ld;
    if (a.alpha != 0 && old > 1) {
        if (a.has_training) {
            for (int i = 0; i < old; ++i) {
                double nf = a.n_f[(int64_t)d * a.nb + i];
                avg[i] = nf > 0 ? a.sum_f[(int64_t)d * a.nb + i] / nf : 0.;
            }
        }
        if (a.alpha > 0) {
            tmp[0] = fabs(7. * avg[0] + avg[1]) / 8.;
            tmp[old - 1] = fabs(7. * avg[old - 1] + avg[old - 2]) / 8.;
            double s = tmp[0] + tmp[old - 1];
            for (int i = 1; i < old - 1; ++i) {
                tmp[i] = fabs(6. * avg[i] + avg[i - 1] + avg[i + 1]) / 8.;
                s += tmp[i];
            }
            if (s > 0) {
                for (int i = 0; i < old; ++i) avg[i] = tmp[i] / s + kTiny;
            } else {
                for (int i = 0; i < old; ++i) avg[i] = kTiny;
            }
            for (int i = 0; i < old; ++i)
                if (avg[i] > 0 && avg[i] <= 0.99999999)
                    avg[i] = pow(-(1 - avg[i]) / log(avg[i]), a.alpha);
        }
    }
    ng[0] = g[0];
    ng[nn] = g[old];
    double acc = 0., f_ninc = 0.;
    int j = -1;
    for (int i = 0; i < old; ++i) f_ninc += avg[i];
    f_ninc /= (double)nn;
    int i = 1;
    for (; i < nn; ++i) {
        bool stopped = false;
        while (acc < f_ninc) {
            j += 1;
            if (j < old) acc += avg[j];
            else { stopped = true; break; }
        }
        if (stopped) break;
        acc -= f_ninc;
        ng[i] = g[j + 1] - (acc / avg[j]) * (g[j + 1] - g[j]);
    }
    // The reference leaves nodes after an early stop unset (round-off only);
    // pin them to the upper limit.
    for (; i < nn; ++i) ng[i] = g[old];
    pad_row(ng, nn, a.new_ld);
}

__global__ void adapt_exact_kernel(AdaptArgs a) {
    if (a.status && *a.status) return;
    if (a.chain) {
        for (int d = 0; d < a.dim; ++d) source_row(a, d, blockDim.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            double* avg = a.buf;
            double* tmp = a.buf + a.nb;
            for (int i = 0; i < a.nb; ++i) avg[i] = 1.0;
            for (int d = 0; d < a.dim; ++d)
                adapt_axis_exact(a, d, avg, tmp, a.old_rows ? a.old_rows + (int64_t)d * a.ld
                                                            : a.grid + (int64_t)d * a.ld);
        }
        __syncthreads();
        for (int d = 0; d < a.dim; ++d) clear_row(a, d, blockDim.x);
    } else {
        const int d = blockIdx.x;
        const double* g = source_row(a, d, blockDim.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            double* avg = a.buf + (int64_t)d * 2 * a.nb;
            double* tmp = avg + a.nb;
            for (int i = 0; i < a.nb; ++i) avg[i] = 1.0;
            adapt_axis_exact(a, d, avg, tmp, g);
        }
        __syncthreads();
        clear_row(a, d, blockDim.x);
    }
}

__global__ __launch_bounds__(1024) void adapt_parallel_kernel(AdaptArgs a) {
    __shared__ double av[kAdaptLdsMax];
    __shared__ double P[kAdaptLdsMax];
    __shared__ double red[40];
    if (a.status && *a.status) return;
    adapt_parallel_axis(a, blockIdx.x, av, P, red);
}

// The integrator's adapt: in place on `grid`, skipped when *status != 0.
// The training data are left as they are (the next iteration's allocation
// kernel clears them).
int adapt_in_place_setup(vgx_ctx* ctx, int dim, const int32_t* ninc, double* grid, int ld, double* sum_f,
                         double* n_f, int has_training, double alpha, int exact, const int32_t* status,
                         AdaptArgs* out, int* parallel) {
    AdaptArgs& a = *out;
    a = AdaptArgs{};
    a.dim = dim;
    a.ld = ld;
    a.nb = ld - 1;
    a.new_ld = ld;
    a.has_training = has_training;
    a.alpha = alpha;
    a.grid = grid;
    a.sum_f = sum_f;
    a.n_f = n_f;
    a.new_grid = grid;
    a.status = status;
    a.clear_sum_f = nullptr;
    a.clear_n_f = nullptr;
    int max_old = 0;
    for (int d = 0; d < dim; ++d) {
        a.ninc[d] = ninc[d];
        a.new_ninc[d] = ninc[d];
        max_old = ninc[d] > max_old ? ninc[d] : max_old;
    }
    a.chain = (!has_training && alpha > 0) ? 1 : 0;
    if (!exact && max_old > kAdaptLdsMax) exact = 1;
    size_t rows = (size_t)dim * ld;
    size_t bufs = (size_t)(a.chain ? 1 : dim) * 2 * (size_t)a.nb;
    int rc = ctx->adapt_tmp.ensure((rows + bufs) * sizeof(double));
    if (rc) return rc;
    a.old_rows = (double*)ctx->adapt_tmp.ptr;
    a.buf = a.old_rows + rows;
    *parallel = !(exact || a.chain);
    return VGX_OK;
}

int adapt_in_place(vgx_ctx* ctx, int dim, const int32_t* ninc, double* grid, int ld, double* sum_f,
                   double* n_f, int has_training, double alpha, int exact, const int32_t* status,
                   hipStream_t st) {
    AdaptArgs a;
    int parallel = 0;
    int rc = adapt_in_place_setup(ctx, dim, ninc, grid, ld, sum_f, n_f, has_training, alpha, exact, status,
                                  &a, &parallel);
    if (rc) return rc;
    if (parallel)
        hipLaunchKernelGGL(adapt_parallel_kernel, dim3(dim), dim3(1024), 0, st, a);
    else
        hipLaunchKernelGGL(adapt_exact_kernel, dim3(a.chain ? 1 : dim), dim3(64), 0, st, a);
    return check_launch();
}

}  // namespace vgx

using namespace vgx;

extern "C" int vgx_adapt(vgx_ctx* ctx, int dim, const int32_t* ninc, const double* grid, int ld,
                         int has_training, const double* sum_f, const double* n_f, double alpha,
                         const int32_t* new_ninc, double* new_grid, int new_ld, int exact,
                         void* stream) {
    if (dim < 1 || dim > VGX_MAXDIM || !ninc || !new_ninc || !grid || !new_grid) return VGX_EINVAL;
    if (grid == new_grid) return VGX_EINVAL;
    if (has_training && (!sum_f || !n_f)) return VGX_EINVAL;
    AdaptArgs a{};
    a.dim = dim;
    a.ld = ld;
    a.nb = ld - 1;
    a.new_ld = new_ld;
    a.has_training = has_training;
    a.alpha = alpha;
    a.grid = grid;
    a.sum_f = sum_f;
    a.n_f = n_f;
    a.new_grid = new_grid;
    int max_new = 0, min_new = 1 << 30, max_old = 0;
    for (int d = 0; d < dim; ++d) {
        if (ninc[d] < 1 || ninc[d] + 1 > ld) return VGX_EINVAL;
        a.ninc[d] = ninc[d];
        a.new_ninc[d] = new_ninc[d];
        max_new = new_ninc[d] > max_new ? new_ninc[d] : max_new;
        min_new = new_ninc[d] < min_new ? new_ninc[d] : min_new;
        max_old = ninc[d] > max_old ? ninc[d] : max_old;
    }
    if (min_new < 1) return VGX_EINVAL;
    if (max_new + 1 > new_ld) return VGX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    // pyx:519-530: every axis collapses to its end points
    if (max_new == 1) {
        a.alpha = 0.0;  // the CDF walk with one increment writes only the ends
        exact = 1;
    }
    // pyx:534: with alpha > 0 and no training data the shared buffer carries
    // one axis' smoothed weights into the next; only the chained walk
    // reproduces that, and only when axes differ in ninc does it matter.
    a.chain = (!has_training && alpha > 0) ? 1 : 0;
    if (!exact && max_old > kAdaptLdsMax) exact = 1;
    if (exact || a.chain) {
        if (!ctx) return VGX_EINVAL;
        size_t need = (size_t)(a.chain ? 1 : dim) * 2 * (size_t)a.nb * sizeof(double);
        int rc = ctx->adapt_tmp.ensure(need);
        if (rc) return rc;
        a.buf = (double*)ctx->adapt_tmp.ptr;
        hipLaunchKernelGGL(adapt_exact_kernel, dim3(a.chain ? 1 : dim), dim3(64), 0, st, a);
    } else {
        hipLaunchKernelGGL(adapt_parallel_kernel, dim3(dim), dim3(1024), 0, st, a);
    }
    return check_launch();
}
