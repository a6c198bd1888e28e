// vgx_cov.hip -- per-hcube sums of multi-component integrands (fcn_size > 1):
// the mean vector and the covariance (or the variances) of one iteration,
// src/vegas/_vegas.pyx:2136-2183 with correlate_integrals on or off.
//
// Work item: (hcube h, component pair q).  A thread owns one pair and a
// strided set of the batch's hcubes; for each hcube it makes the reference's
// two passes over the hcube's points: S = sum w f (per component), then the
// deviations d = w f - S/n with the EPSILON guard (a guarded deviation adds
// (EPSILON |S/n|)^2 to its own diagonal entry and is 0 in every other sum).
// Diagonal pairs also contribute S + D to the mean (Neely's correction).
// Sums over hcubes are double-double, reduced in a fixed order (partials per
// workgroup, then one thread per pair), so results do not depend on timing.
#include "vgx_internal.h"

namespace vgx {

constexpr int kCovThreads = 256;
constexpr int kCovMaxBlocks = 256;  // workgroups per pair

// Pair q of the lower triangle (t <= s), row-major: (0,0) (1,0) (1,1) (2,0) ...
__device__ inline void cov_pair(int q, int correlate, int& s, int& t) {
    if (!correlate) {
        s = t = q;
        return;
    }
    int r = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= q) ++r;
    while (r * (r + 1) / 2 > q) --r;
    s = r;
    t = q - r * (r + 1) / 2;
}

__global__ __launch_bounds__(kCovThreads) void cov_hcube_kernel(const uint32_t* off, uint32_t h0, uint32_t h1,
                                                                 uint32_t p0, uint32_t p1, const double* w,
                                                                 const double* f, int m, int correlate,
                                                                 double* part, int32_t* status) {
    __shared__ double red[4 * (kCovThreads / 64)];
    int s, t;
    cov_pair((int)blockIdx.y, correlate, s, t);
    const double eps = kEpsilon;
    DD am{0, 0}, av{0, 0};
    for (uint32_t h = h0 + blockIdx.x * blockDim.x + threadIdx.x; h < h1; h += gridDim.x * blockDim.x) {
        const uint32_t oa = off[h], ob = off[h + 1];
        if (oa < p0 || ob > p1 || ob < oa + 2) {  // not this batch's points: nothing is read
            if (status) atomicOr(status, VGX_ERANGE);
            continue;
        }
        const uint32_t a = oa - p0, b = ob - p0;
        const double n = (double)(b - a);
        double Ss = 0.0, St = 0.0;
        for (uint32_t j = a; j < b; ++j) Ss += w[j] * f[(size_t)j * m + s];
        if (s != t)
            for (uint32_t j = a; j < b; ++j) St += w[j] * f[(size_t)j * m + t];
        const double ms = Ss / n, mt = St / n;
        const double gs = eps * fabs(ms), gt = eps * fabs(mt);
        double Q = 0.0, Ds = 0.0, Dt = 0.0;
        for (uint32_t j = a; j < b; ++j) {
            double ds = w[j] * f[(size_t)j * m + s] - ms;
            if (s == t) {
                if (fabs(ds) < gs) {
                    Q += gs * gs;
                    ds = 0.0;
                } else {
                    Q += ds * ds;
                }
                Ds += ds;
            } else {
                double dt = w[j] * f[(size_t)j * m + t] - mt;
                if (fabs(ds) < gs) ds = 0.0;
                if (fabs(dt) < gt) dt = 0.0;
                Q += ds * dt;
                Ds += ds;
                Dt += dt;
            }
        }
        if (s == t) {
            dd_add(am, Ss + Ds);
            Dt = Ds;
        }
        dd_add(av, (n * Q - Ds * Dt) / (n - 1.));
    }
    am = wave_sum_dd(am);
    av = wave_sum_dd(av);
    const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[4 * wv + 0] = am.hi;
        red[4 * wv + 1] = am.lo;
        red[4 * wv + 2] = av.hi;
        red[4 * wv + 3] = av.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD m2{0, 0}, v2{0, 0};
        for (int k = 0; k < nw; ++k) {
            m2 = dd_sum(m2, DD{red[4 * k], red[4 * k + 1]});
            v2 = dd_sum(v2, DD{red[4 * k + 2], red[4 * k + 3]});
        }
        double* o = part + 4 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
        o[0] = m2.hi;
        o[1] = m2.lo;
        o[2] = v2.hi;
        o[3] = v2.lo;
    }
}

// acc: mean[m] then cov entries [npairs], each a DD (hi, lo) pair; this
// batch's partials are added in workgroup order.
__global__ void cov_reduce_kernel(const double* part, int nbx, int npairs, int m, int correlate, double* acc) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    int s, t;
    cov_pair(q, correlate, s, t);
    DD mm{0, 0}, vv{0, 0};
    for (int b = 0; b < nbx; ++b) {
        const double* p = part + 4 * ((size_t)q * nbx + b);
        mm = dd_sum(mm, DD{p[0], p[1]});
        vv = dd_sum(vv, DD{p[2], p[3]});
    }
    if (s == t) {
        const DD r = dd_sum(DD{acc[2 * s], acc[2 * s + 1]}, mm);
        acc[2 * s] = r.hi;
        acc[2 * s + 1] = r.lo;
    }
    double* cv = acc + 2 * m;
    const DD r = dd_sum(DD{cv[2 * q], cv[2 * q + 1]}, vv);
    cv[2 * q] = r.hi;
    cv[2 * q + 1] = r.lo;
}

}  // namespace vgx

using namespace vgx;

extern "C" int vgx_iter_accumulate_cov(vgx_ctx* ctx, const vgx_config* c, int64_t h0, int64_t h1, int64_t p0,
                                       int64_t p1, const double* f, const double* wgt, int m, int correlate,
                                       double* acc, int32_t* status, void* stream) {
    if (!ctx || !c || !f || !wgt || !acc || m < 1 || m > VGX_MAXCOMP) return VGX_EINVAL;
    if (h0 < 0 || h1 > c->nhcube || h0 >= h1 || p0 < 0 || p1 <= p0 || p1 > 0xFFFFFFFFll) return VGX_EINVAL;
    if (!ctx->off.ptr || ctx->off.bytes < (size_t)(h1 + 1) * sizeof(uint32_t)) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    const int npairs = correlate ? m * (m + 1) / 2 : m;
    const int64_t nh = h1 - h0;
    const int nbx = (int)std::min<int64_t>(kCovMaxBlocks, (nh + kCovThreads - 1) / kCovThreads);
    int rc;
    if ((rc = ctx->cov_part.ensure((size_t)npairs * nbx * 4 * sizeof(double)))) return rc;
    hipLaunchKernelGGL(cov_hcube_kernel, dim3(nbx, npairs), dim3(kCovThreads), 0, s,
                       (const uint32_t*)ctx->off.ptr, (uint32_t)h0, (uint32_t)h1, (uint32_t)p0, (uint32_t)p1, wgt,
                       f, m, correlate, (double*)ctx->cov_part.ptr, status);
    hipLaunchKernelGGL(cov_reduce_kernel, dim3((npairs + 63) / 64), dim3(64), 0, s,
                       (const double*)ctx->cov_part.ptr, nbx, npairs, m, correlate, acc);
    return check_launch();
}
