// vgx_map.hip -- AdaptiveMap.map / invmap / jac1d on gfx950.
//
// The per-axis grid (FP64, dim x ld) is staged in LDS once per workgroup;
// each thread maps whole points (grid-stride loop).  Increments are formed
// as grid[k+1]-grid[k], the very subtraction the reference stores in `inc`
// (src/vegas/_vegas.pyx:130-131, 236-237, 589-592), so they are bit-equal.
#include "vgx_internal.h"

namespace vgx {

struct MapArgs {
    int dim;
    int ld;
    int use_lds;
    int pad;
    int32_t ninc[VGX_MAXDIM];
    const double* grid;
    int64_t n;
};

__device__ inline const double* stage_grid(const MapArgs& a, double* lds) {
    if (!a.use_lds) return a.grid;
    const int tot = a.dim * a.ld;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) lds[i] = a.grid[i];
    __syncthreads();
    return lds;
}

// One axis of AdaptiveMap.map (pyx:347-359).  Returns x, multiplies J.
__device__ inline double map_axis(const double* g, int ni, double yv, double& J, int* kout) {
    double t = yv * (double)ni;
    t = fmin(t, (double)ni);  // y > 1 takes the k >= ninc branch, as in the reference
    int k = (int)floor(t);
    k = k < 0 ? 0 : k;        // y < 0 is never produced; keep the read in bounds
    double dl = t - (double)k;
    double xv;
    if (k < ni) {
        double inc = g[k + 1] - g[k];
        xv = g[k] + inc * dl;
        J *= inc * (double)ni;
    } else {
        xv = g[ni];
        J *= (g[ni] - g[ni - 1]) * (double)ni;
    }
    *kout = k;
    return xv;
}

template <typename T>
__global__ __launch_bounds__(256) void map_kernel(MapArgs a, const T* __restrict__ y,
                                                  T* __restrict__ x, T* __restrict__ jac) {
    extern __shared__ double lds[];
    const double* G = stage_grid(a, lds);
    const int dim = a.dim;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double J = 1.0;
        for (int d = 0; d < dim; ++d) {
            int k;
            double xv = map_axis(G + d * a.ld, a.ninc[d], (double)y[i * dim + d], J, &k);
            x[i * dim + d] = (T)xv;
        }
        jac[i] = (T)J;
    }
}

// AdaptiveMap.invmap (pyx:362-416): searchsorted(side='right') over the
// live nodes of each axis, then the three branches of the reference.
__global__ __launch_bounds__(256) void invmap_kernel(MapArgs a, const double* __restrict__ x,
                                                     double* __restrict__ y,
                                                     double* __restrict__ jac) {
    extern __shared__ double lds[];
    const double* G = stage_grid(a, lds);
    const int dim = a.dim;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double J = 1.0;
        for (int d = 0; d < dim; ++d) {
            const double* g = G + d * a.ld;
            const int ni = a.ninc[d];
            const double xv = x[i * dim + d];
            int lo = 0, hi = ni + 1;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (g[mid] <= xv) lo = mid + 1; else hi = mid;
            }
            const int iy = lo;
            double yv;
            if (iy > 0 && iy <= ni) {
                const int k = iy - 1;
                const double inc = g[k + 1] - g[k];
                yv = ((double)k + (xv - g[k]) / inc) / (double)ni;
                J *= inc * (double)ni;
            } else if (iy <= 0) {
                yv = 0.;
                J *= (g[1] - g[0]) * (double)ni;
            } else {
                yv = 1.0;
                J *= (g[ni] - g[ni - 1]) * (double)ni;
            }
            y[i * dim + d] = yv;
        }
        jac[i] = J;
    }
}

// AdaptiveMap.jac1d (pyx:265-295)
__global__ __launch_bounds__(256) void jac1d_kernel(MapArgs a, const double* __restrict__ y,
                                                    double* __restrict__ jac) {
    extern __shared__ double lds[];
    const double* G = stage_grid(a, lds);
    const int dim = a.dim;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int d = 0; d < dim; ++d) {
            double J = 1.0;
            int k;
            map_axis(G + d * a.ld, a.ninc[d], y[i * dim + d], J, &k);
            jac[i * dim + d] = J;
        }
    }
}

static int fill_map_args(MapArgs& a, int dim, const int32_t* ninc, const double* grid, int ld,
                         int64_t n) {
    if (dim < 1 || dim > VGX_MAXDIM || n < 0 || !ninc || !grid) return VGX_EINVAL;
    a.dim = dim;
    a.ld = ld;
    a.grid = grid;
    a.n = n;
    for (int d = 0; d < dim; ++d) {
        if (ninc[d] < 1 || ninc[d] + 1 > ld) return VGX_EINVAL;
        a.ninc[d] = ninc[d];
    }
    const size_t bytes = (size_t)dim * ld * sizeof(double);
    a.use_lds = bytes <= 64 * 1024;
    return VGX_OK;
}

static unsigned map_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace vgx

using namespace vgx;

extern "C" int vgx_map_f64(int dim, const int32_t* ninc, const double* grid, int ld,
                           const double* y, double* x, double* jac, int64_t n, void* stream) {
    MapArgs a;
    int rc = fill_map_args(a, dim, ninc, grid, ld, n);
    if (rc) return rc;
    if (n == 0) return VGX_OK;
    size_t sh = a.use_lds ? (size_t)dim * ld * sizeof(double) : 0;
    hipLaunchKernelGGL(map_kernel<double>, dim3(map_blocks(n)), dim3(256), sh,
                       (hipStream_t)stream, a, y, x, jac);
    return check_launch();
}

extern "C" int vgx_map_f32(int dim, const int32_t* ninc, const double* grid, int ld,
                           const float* y, float* x, float* jac, int64_t n, void* stream) {
    MapArgs a;
    int rc = fill_map_args(a, dim, ninc, grid, ld, n);
    if (rc) return rc;
    if (n == 0) return VGX_OK;
    size_t sh = a.use_lds ? (size_t)dim * ld * sizeof(double) : 0;
    hipLaunchKernelGGL(map_kernel<float>, dim3(map_blocks(n)), dim3(256), sh,
                       (hipStream_t)stream, a, y, x, jac);
    return check_launch();
}

extern "C" int vgx_invmap_f64(int dim, const int32_t* ninc, const double* grid, int ld,
                              const double* x, double* y, double* jac, int64_t n, void* stream) {
    MapArgs a;
    int rc = fill_map_args(a, dim, ninc, grid, ld, n);
    if (rc) return rc;
    if (n == 0) return VGX_OK;
    size_t sh = a.use_lds ? (size_t)dim * ld * sizeof(double) : 0;
    hipLaunchKernelGGL(invmap_kernel, dim3(map_blocks(n)), dim3(256), sh, (hipStream_t)stream,
                       a, x, y, jac);
    return check_launch();
}

extern "C" int vgx_jac1d_f64(int dim, const int32_t* ninc, const double* grid, int ld,
                             const double* y, double* jac, int64_t n, void* stream) {
    MapArgs a;
    int rc = fill_map_args(a, dim, ninc, grid, ld, n);
    if (rc) return rc;
    if (n == 0) return VGX_OK;
    size_t sh = a.use_lds ? (size_t)dim * ld * sizeof(double) : 0;
    hipLaunchKernelGGL(jac1d_kernel, dim3(map_blocks(n)), dim3(256), sh, (hipStream_t)stream, a,
                       y, jac);
    return check_launch();
}
