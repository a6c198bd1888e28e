// vgx_cov.hip -- per-hcube sums of multi-component integrands (fcn_size > 1):
// the mean vector and the covariance (or the variances) of one iteration,
// src/vegas/_vegas.pyx:2136-2183 with correlate_integrals on or off.
//
// Work item: (hcube h, component pair q, or all pairs for m <= 4).  A thread
// owns a strided set of the batch's hcubes (the long ones are shared by its
// workgroup); for each hcube it makes the reference's two passes over the
// hcube's points: S = sum w f (per component), then the deviations
// d = w f - S/n with the EPSILON guard (a guarded deviation adds
// (EPSILON |S/n|)^2 to its own diagonal entry and is 0 in every other sum).
// Diagonal pairs also contribute S + D to the mean (Neely's correction).
// Sums over hcubes are double-double, reduced in a fixed order (partials per
// workgroup, then one thread per pair), so results do not depend on timing.
#include "vgx_internal.h"

namespace vgx {

constexpr int kCovThreads = 256;
constexpr int kCovWaves = kCovThreads / 64;
constexpr int kCovMaxBlocks = 8192;  // workgroups per pair (one thread per hcube up to 2M hcubes)
constexpr uint32_t kCovLong = 32;    // hcubes with more points are summed by the whole workgroup

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

// Workgroup sum of K values per thread, in a fixed order (wave butterflies,
// then waves 0..3); every thread receives the totals.  red: K * kCovWaves.
template <int K>
__device__ inline void block_sum(double (&v)[K], double* red) {
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double r = wave_sum_fixed(v[k]);
        if ((threadIdx.x & 63) == 0) red[wv * K + k] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double r = red[k];
#pragma unroll
        for (int u = 1; u < kCovWaves; ++u) r += red[u * K + k];
        v[k] = r;
    }
    __syncthreads();
}

// The hcubes [h0, h1) in rounds of gridDim.x * blockDim.x: thread t of
// workgroup b takes hcube h0 + r0 + t * gridDim.x + b (neighbouring hcubes,
// which tend to be large together near a peak, land in different
// workgroups).  A hcube of at most kCovLong points is summed by its thread
// (wk.thread_hcube); longer ones are summed after the round by the whole
// workgroup, in thread order (wk.block_hcube), so the work per thread stays
// bounded however the points are spread.  A hcube whose points are not the
// batch's [p0, p1) (or that has fewer than two) sets VGX_ERANGE and is not read.
template <class Work>
__device__ inline void cov_rounds(const uint32_t* off, uint32_t h0, uint32_t h1, uint32_t p0, uint32_t p1,
                                  int32_t* status, Work& wk) {
    __shared__ uint64_t lmask[kCovWaves];
    __shared__ uint32_t la[kCovThreads], lb[kCovThreads];
    const uint32_t nh = h1 - h0, stride = gridDim.x * blockDim.x;
    for (uint32_t r0 = 0; r0 < nh; r0 += stride) {
        const uint32_t idx = r0 + threadIdx.x * gridDim.x + blockIdx.x;
        bool lng = false;
        if (idx < nh) {
            const uint32_t oa = off[h0 + idx], ob = off[h0 + idx + 1];
            if (oa < p0 || ob > p1 || ob < oa + 2) {
                if (status) atomicOr(status, VGX_ERANGE);
            } else if (ob - oa > kCovLong) {
                lng = true;
                la[threadIdx.x] = oa - p0;
                lb[threadIdx.x] = ob - p0;
            } else {
                wk.thread_hcube(oa - p0, ob - p0);
            }
        }
        const uint64_t bal = __ballot(lng);
        if ((threadIdx.x & 63) == 0) lmask[threadIdx.x >> 6] = bal;
        __syncthreads();
        for (int wv = 0; wv < kCovWaves; ++wv) {
            uint64_t mk = lmask[wv];
            while (mk) {
                const int src = wv * 64 + __builtin_ctzll(mk);
                mk &= mk - 1;
                wk.block_hcube(la[src], lb[src]);
            }
        }
        __syncthreads();
    }
}

// One component pair (s, t) per workgroup row blockIdx.y, any m.
struct PairWork {
    const double* w;
    const double* f;
    int m, s, t;
    double eps;
    DD am, av;
    double* red;

    // the reference's two passes over one hcube's points, [a, b)
    __device__ void thread_hcube(uint32_t a, uint32_t b) {
        const double n = (double)(b - a);
        double Ss = 0.0, St = 0.0;
        for (uint32_t j = a; j < b; ++j) Ss += w[j] * f[(size_t)j * m + s];
        if (s != t)
            for (uint32_t j = a; j < b; ++j) St += w[j] * f[(size_t)j * m + t];
        double v[3] = {0.0, 0.0, 0.0};
        pass2(a, b, 1, Ss / n, St / n, v);
        close(n, Ss, v);
    }
    __device__ void block_hcube(uint32_t a, uint32_t b) {
        const double n = (double)(b - a);
        double S[2] = {0.0, 0.0};
        for (uint32_t j = a + threadIdx.x; j < b; j += blockDim.x) {
            S[0] += w[j] * f[(size_t)j * m + s];
            S[1] += w[j] * f[(size_t)j * m + t];
        }
        block_sum<2>(S, red);
        double v[3] = {0.0, 0.0, 0.0};
        pass2(a + threadIdx.x, b, blockDim.x, S[0] / n, S[1] / n, v);
        block_sum<3>(v, red);
        if (threadIdx.x == 0) close(n, S[0], v);
    }
    // v = {Q, Ds, Dt} over points a, a + step, ... < b
    __device__ void pass2(uint32_t a, uint32_t b, uint32_t step, double ms, double mt, double (&v)[3]) {
        const double gs = eps * fabs(ms), gt = eps * fabs(mt);
        for (uint32_t j = a; j < b; j += step) {
            double ds = w[j] * f[(size_t)j * m + s] - ms;
            if (s == t) {
                if (fabs(ds) < gs) {
                    v[0] += gs * gs;
                    ds = 0.0;
                } else {
                    v[0] += ds * ds;
                }
                v[1] += ds;
            } else {
                double dt = w[j] * f[(size_t)j * m + t] - mt;
                if (fabs(ds) < gs) ds = 0.0;
                if (fabs(dt) < gt) dt = 0.0;
                v[0] += ds * dt;
                v[1] += ds;
                v[2] += dt;
            }
        }
    }
    __device__ void close(double n, double Ss, const double (&v)[3]) {
        double Dt = v[2];
        if (s == t) {
            dd_add(am, Ss + v[1]);
            Dt = v[1];
        }
        dd_add(av, (n * v[0] - v[1] * Dt) / (n - 1.));
    }
};

__global__ __launch_bounds__(kCovThreads) void cov_hcube_kernel(const uint32_t* off, uint32_t h0, uint32_t h1,
                                                                 uint32_t p0, uint32_t p1, const double* w,
                                                                 const double* f, int m, int correlate,
                                                                 double* part, int32_t* status) {
    __shared__ double red[4 * kCovWaves];
    PairWork wk{w, f, m, 0, 0, kEpsilon, DD{0, 0}, DD{0, 0}, red};
    cov_pair((int)blockIdx.y, correlate, wk.s, wk.t);
    cov_rounds(off, h0, h1, p0, p1, status, wk);
    DD am = wave_sum_dd(wk.am);
    DD av = wave_sum_dd(wk.av);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[4 * wv + 0] = am.hi;
        red[4 * wv + 1] = am.lo;
        red[4 * wv + 2] = av.hi;
        red[4 * wv + 3] = av.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD m2{0, 0}, v2{0, 0};
        for (int k = 0; k < kCovWaves; ++k) {
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

// Up to kCovAllMax components: every pair at once from the hcube's
// contiguous rows of f (each row read once per pass), plain FP64 sums per
// thread, then the same fixed-order double-double reduction and partial
// layout as cov_hcube_kernel.
constexpr int kCovAllMax = 4;

template <int M>
struct AllWork {
    static constexpr int NP = M * (M + 1) / 2;
    const double* w;
    const double* f;
    double eps;
    double am[M], av[NP];
    double* red;

    __device__ void thread_hcube(uint32_t a, uint32_t b) {
        double S[M];
#pragma unroll
        for (int s = 0; s < M; ++s) S[s] = 0.0;
        pass1(a, b, 1, S);
        double v[M + NP];
        pass2(a, b, 1, S, (double)(b - a), v);
        close((double)(b - a), S, v);
    }
    __device__ void block_hcube(uint32_t a, uint32_t b) {
        double S[M];
#pragma unroll
        for (int s = 0; s < M; ++s) S[s] = 0.0;
        pass1(a + threadIdx.x, b, blockDim.x, S);
        block_sum<M>(S, red);
        double v[M + NP];
        pass2(a + threadIdx.x, b, blockDim.x, S, (double)(b - a), v);
        block_sum<M + NP>(v, red);
        if (threadIdx.x == 0) close((double)(b - a), S, v);
    }
    __device__ void pass1(uint32_t a, uint32_t b, uint32_t step, double (&S)[M]) {
        for (uint32_t j = a; j < b; j += step) {
            const double wj = w[j];
#pragma unroll
            for (int s = 0; s < M; ++s) S[s] += wj * f[(size_t)j * M + s];
        }
    }
    // v = {D[M], Q[NP]} over points a, a + step, ... < b; S = the hcube's sums
    __device__ void pass2(uint32_t a, uint32_t b, uint32_t step, const double (&S)[M], double n,
                          double (&v)[M + NP]) {
        double mu[M], g[M];
#pragma unroll
        for (int s = 0; s < M; ++s) {
            mu[s] = S[s] / n;
            g[s] = eps * fabs(mu[s]);
        }
#pragma unroll
        for (int k = 0; k < M + NP; ++k) v[k] = 0.0;
        for (uint32_t j = a; j < b; j += step) {
            const double wj = w[j];
            double d[M];
#pragma unroll
            for (int s = 0; s < M; ++s) {
                d[s] = wj * f[(size_t)j * M + s] - mu[s];
                double& qss = v[M + s * (s + 1) / 2 + s];
                if (fabs(d[s]) < g[s]) {
                    qss += g[s] * g[s];
                    d[s] = 0.0;
                } else {
                    qss += d[s] * d[s];
                }
                v[s] += d[s];
            }
#pragma unroll
            for (int s = 1; s < M; ++s)
#pragma unroll
                for (int t = 0; t < s; ++t) v[M + s * (s + 1) / 2 + t] += d[s] * d[t];
        }
    }
    __device__ void close(double n, const double (&S)[M], const double (&v)[M + NP]) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            am[s] += S[s] + v[s];
#pragma unroll
            for (int t = 0; t <= s; ++t) {
                const int q = s * (s + 1) / 2 + t;
                av[q] += (n * v[M + q] - v[s] * v[t]) / (n - 1.);
            }
        }
    }
};

template <int M>
__global__ __launch_bounds__(kCovThreads) void cov_hcube_all_kernel(const uint32_t* off, uint32_t h0, uint32_t h1,
                                                                     uint32_t p0, uint32_t p1, const double* w,
                                                                     const double* f, int correlate,
                                                                     double* part, int32_t* status) {
    constexpr int NP = M * (M + 1) / 2;
    __shared__ double red[kCovWaves * 2 * (M + NP)];
    AllWork<M> wk;
    wk.w = w;
    wk.f = f;
    wk.eps = kEpsilon;
    wk.red = red;
#pragma unroll
    for (int s = 0; s < M; ++s) wk.am[s] = 0.0;
#pragma unroll
    for (int q = 0; q < NP; ++q) wk.av[q] = 0.0;
    cov_rounds(off, h0, h1, p0, p1, status, wk);
    // workgroup partials: mean entries then pair entries, DD per wave
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < M + NP; ++k) {
        const DD r = wave_sum_dd(DD{k < M ? wk.am[k] : wk.av[k - M], 0.0});
        if ((threadIdx.x & 63) == 0) {
            red[2 * ((size_t)wv * (M + NP) + k)] = r.hi;
            red[2 * ((size_t)wv * (M + NP) + k) + 1] = r.lo;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)(correlate ? NP : M)) {
        // output pair q (the layout of cov_hcube_kernel: npairs x blocks x 4)
        const int q = threadIdx.x;
        int s, t;
        cov_pair(q, correlate, s, t);
        const int kq = s * (s + 1) / 2 + t;
        DD m2{0, 0}, v2{0, 0};
        for (int k = 0; k < kCovWaves; ++k) {
            if (s == t) m2 = dd_sum(m2, DD{red[2 * ((size_t)k * (M + NP) + s)], red[2 * ((size_t)k * (M + NP) + s) + 1]});
            v2 = dd_sum(v2, DD{red[2 * ((size_t)k * (M + NP) + M + kq)], red[2 * ((size_t)k * (M + NP) + M + kq) + 1]});
        }
        double* o = part + 4 * ((size_t)q * gridDim.x + blockIdx.x);
        o[0] = m2.hi;
        o[1] = m2.lo;
        o[2] = v2.hi;
        o[3] = v2.lo;
    }
}

// acc: mean[m] then cov entries [npairs], each a DD (hi, lo) pair.  One
// workgroup per pair sums that pair's workgroup partials (thread k takes
// partials k, k + 256, ..., then a fixed tree) and adds them to acc.
__global__ __launch_bounds__(kCovThreads) void cov_reduce_kernel(const double* part, int nbx, int npairs, int m,
                                                                 int correlate, double* acc) {
    __shared__ double red[4 * (kCovThreads / 64)];
    const int q = blockIdx.x;
    int s, t;
    cov_pair(q, correlate, s, t);
    DD mm{0, 0}, vv{0, 0};
    for (int b = threadIdx.x; b < nbx; b += blockDim.x) {
        const double* p = part + 4 * ((size_t)q * nbx + b);
        mm = dd_sum(mm, DD{p[0], p[1]});
        vv = dd_sum(vv, DD{p[2], p[3]});
    }
    mm = wave_sum_dd(mm);
    vv = wave_sum_dd(vv);
    const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[4 * wv] = mm.hi;
        red[4 * wv + 1] = mm.lo;
        red[4 * wv + 2] = vv.hi;
        red[4 * wv + 3] = vv.lo;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    mm = DD{0, 0};
    vv = DD{0, 0};
    for (int k = 0; k < nw; ++k) {
        mm = dd_sum(mm, DD{red[4 * k], red[4 * k + 1]});
        vv = dd_sum(vv, DD{red[4 * k + 2], red[4 * k + 3]});
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
    const uint32_t* off = (const uint32_t*)ctx->off.ptr;
    double* part = (double*)ctx->cov_part.ptr;
    // timed as one span: per-hcube sums and their reduction (events around both)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = timed_slot(ctx, kTimedCov, &e0, &e1);
    if (timed) VGX_HIP(hipEventRecord(e0, s));
    switch (m <= kCovAllMax ? m : 0) {
        case 1: hipLaunchKernelGGL(cov_hcube_all_kernel<1>, dim3(nbx), dim3(kCovThreads), 0, s, off, (uint32_t)h0,
                                   (uint32_t)h1, (uint32_t)p0, (uint32_t)p1, wgt, f, correlate, part, status); break;
        case 2: hipLaunchKernelGGL(cov_hcube_all_kernel<2>, dim3(nbx), dim3(kCovThreads), 0, s, off, (uint32_t)h0,
                                   (uint32_t)h1, (uint32_t)p0, (uint32_t)p1, wgt, f, correlate, part, status); break;
        case 3: hipLaunchKernelGGL(cov_hcube_all_kernel<3>, dim3(nbx), dim3(kCovThreads), 0, s, off, (uint32_t)h0,
                                   (uint32_t)h1, (uint32_t)p0, (uint32_t)p1, wgt, f, correlate, part, status); break;
        case 4: hipLaunchKernelGGL(cov_hcube_all_kernel<4>, dim3(nbx), dim3(kCovThreads), 0, s, off, (uint32_t)h0,
                                   (uint32_t)h1, (uint32_t)p0, (uint32_t)p1, wgt, f, correlate, part, status); break;
        default:
            hipLaunchKernelGGL(cov_hcube_kernel, dim3(nbx, npairs), dim3(kCovThreads), 0, s, off, (uint32_t)h0,
                               (uint32_t)h1, (uint32_t)p0, (uint32_t)p1, wgt, f, m, correlate, part, status);
    }
    hipLaunchKernelGGL(cov_reduce_kernel, dim3(npairs), dim3(kCovThreads), 0, s,
                       (const double*)ctx->cov_part.ptr, nbx, npairs, m, correlate, acc);
    if (timed) VGX_HIP(hipEventRecord(e1, s));
    return check_launch();
}
