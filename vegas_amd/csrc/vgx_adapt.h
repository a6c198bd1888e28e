// vgx_adapt.h -- the block-parallel adapt of one axis (AdaptiveMap.adapt,
// src/vegas/_vegas.pyx:467-594), shared by the single-map and batched kernels.
#pragma once
#include "vgx_internal.h"

namespace vgx {

// Deterministic block-wide sum (fixed tree over lanes, then waves).
// red: at least 33 doubles of LDS.
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum_fixed(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < nw; ++k) t += red[k];
        red[32] = t;
    }
    __syncthreads();
    t = red[32];
    __syncthreads();
    return t;
}

// One axis by one workgroup.  av[0..old): the averaged training data
// sum_f/n_f (or 1 without training), in LDS; g: the old row; ng: the new row
// (new_ld stride, padding repeats the last node); P: LDS scratch of old
// doubles.  alpha > 0: 3-tap smoothing, normalisation and the compression
// (-(1-a)/ln a)^alpha (pyx:535-556); then node i of nn lands in the first old
// increment j whose prefix P_j >= i F, F = P_old / nn (the increment the
// reference's sequential walk, pyx:563-587, stops in).
__device__ inline void adapt_axis_block(double* av, double* P, double* red, const double* g, double* ng,
                                        int old, int nn, double alpha, int new_ld) {
    const int T = blockDim.x;
    if (alpha > 0 && old > 1) {
        double part = 0.0;
        for (int i = threadIdx.x; i < old; i += T) {
            double t;
            if (i == 0) t = fabs(7. * av[0] + av[1]) / 8.;
            else if (i == old - 1) t = fabs(7. * av[old - 1] + av[old - 2]) / 8.;
            else t = fabs(6. * av[i] + av[i - 1] + av[i + 1]) / 8.;
            P[i] = t;
            part += t;
        }
        const double s = block_sum(part, red);
        for (int i = threadIdx.x; i < old; i += T) {
            double v = s > 0 ? P[i] / s + kTiny : kTiny;
            if (v > 0 && v <= 0.99999999) v = pow(-(1 - v) / log(v), alpha);
            av[i] = v;
        }
        __syncthreads();
    }
    // inclusive prefix sums of av, chunk by chunk of T: a DPP scan within
    // each wave, the wave totals through LDS (red[0..T/64)), then each wave's
    // offset = carry + the totals of the waves before it (fixed order; two
    // barriers per chunk)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = T >> 6;
    double carry = 0.0;
    for (int base = 0; base < old; base += T) {
        const int i = base + threadIdx.x;
        double v = wave_segscan(i < old ? av[i] : 0.0, lane == 0);
        if (lane == 63) red[wv] = v;
        __syncthreads();
        double off = carry, tot = carry;
        for (int k = 0; k < nw; ++k) {
            if (k < wv) off += red[k];
            tot += red[k];
        }
        if (i < old) P[i] = off + v;
        carry = tot;
        __syncthreads();
    }
    const double total = P[old - 1];
    const double F = total / (double)nn;
    const double glo = g[0], ghi = g[old];
    for (int i = 1 + threadIdx.x; i < nn; i += T) {
        const double target = (double)i * F;
        int lo = 0, hi = old;  // first j with P[j] >= target
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (P[mid] >= target) hi = mid; else lo = mid + 1;
        }
        double xv;
        if (lo >= old) {
            xv = ghi;
        } else {
            const int j = lo;
            const double acc = P[j] - target;
            xv = g[j + 1] - (acc / av[j]) * (g[j + 1] - g[j]);
        }
        ng[i] = xv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ng[0] = glo;
        ng[nn] = ghi;
        for (int i = nn + 1; i < new_ld; ++i) ng[i] = ng[nn];
    }
}

constexpr int kAdaptLdsMax = 4000;  // old increments handled in LDS (2 x 31.25 KB)

struct AdaptArgs {
    int dim;
    int ld;       // old grid stride
    int nb;       // ld - 1, stride of sum_f / n_f
    int new_ld;
    int has_training;
    int chain;    // exact mode: one lane for all axes, buffer shared (pyx:534)
    double alpha;
    int32_t ninc[VGX_MAXDIM];
    int32_t new_ninc[VGX_MAXDIM];
    const double* grid;
    const double* sum_f;
    const double* n_f;
    double* new_grid;
    double* buf;  // exact mode scratch: avg and tmp, 2 x nb per axis (chain: one pair)
    // in-place use by the integrator (all null for vgx_adapt):
    double* old_rows;      // dim x ld copy of the old grid rows
    double* clear_sum_f;   // training data to clear once read
    double* clear_n_f;
    const int32_t* status; // skip everything when nonzero (NaN seen)
};

// In-place mode: copy the axis' old row aside before it is overwritten and
// return the row to read from.
__device__ inline const double* source_row(const AdaptArgs& a, int d, int lanes) {
    const double* g = a.grid + (int64_t)d * a.ld;
    if (!a.old_rows) return g;
    double* o = a.old_rows + (int64_t)d * a.ld;
    for (int i = threadIdx.x; i < a.ld; i += lanes) o[i] = g[i];
    return o;
}

__device__ inline void clear_row(const AdaptArgs& a, int d, int lanes) {
    if (!a.clear_sum_f) return;
    for (int i = threadIdx.x; i < a.nb; i += lanes) {
        a.clear_sum_f[(int64_t)d * a.nb + i] = 0.0;
        a.clear_n_f[(int64_t)d * a.nb + i] = kTiny;
    }
}

// The parallel adapt of axis d by the calling workgroup, in place when
// a.old_rows is set (av, P: kAdaptLdsMax doubles of LDS each; red: 40).
__device__ inline void adapt_parallel_axis(const AdaptArgs& a, int d, double* av, double* P, double* red) {
    const int old = a.ninc[d], nn = a.new_ninc[d];
    double* ng = a.new_grid + (int64_t)d * a.new_ld;
    const int T = blockDim.x;
    const double* g = source_row(a, d, T);
    for (int i = threadIdx.x; i < old; i += T) {
        double v = 1.0;
        if (a.alpha != 0 && old > 1 && a.has_training) {
            double nf = a.n_f[(int64_t)d * a.nb + i];
            v = nf > 0 ? a.sum_f[(int64_t)d * a.nb + i] / nf : 0.;
        }
        av[i] = v;
    }
    __syncthreads();
    clear_row(a, d, T);
    adapt_axis_block(av, P, red, g, ng, old, nn, a.alpha, a.new_ld);
}

}  // namespace vgx
