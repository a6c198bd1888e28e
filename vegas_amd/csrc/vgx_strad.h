// vgx_strad.h -- the straddler kernel: hcubes cut by wave-tile edges.
#pragma once
#include "vgx_fused.h"

namespace vgx {

// ---------------------------------------------------------- straddlers ----
// Hcubes cut by a tile edge.  For each tile part the fused kernel left
// (n_u, S_u, Q_u, D_u): count, sum and guarded two-pass sums about the part
// mean m_u = S_u/n_u.  The slot of the tile where the hcube starts owns it
// and combines the parts exactly: with m = S/n and d_u = m_u - m,
//   sum (w f - m)^2 = sum_u [Q_u + 2 d_u D_u + n_u d_u^2],
//   sum (w f - m)   = sum_u [D_u + n_u d_u],
// where a shift below the reference's guard EPSILON |m| counts as zero as a
// deviation does (pyx:2157-2163).  Then the per-hcube results
// (pyx:2177-2193).  One lane per slot; owners spanning more than kLaneParts
// tiles are combined by the whole wave.  For u > first tile the part is
// always the tile's first segment (slot 2u).
struct StradArgs {
    double* errv;
    const uint32_t* off;
    const Slot* slots;
    double* sigf;
    double* strad_part;   // [gridDim.x][6]: DD mean, var, sigf sums
    double* cont;         // or, when set: [slot][3] = S + D, var_h, sigf_h of an
                          // owner slot, 0 for the others (batched integrals)
    Scalars* sc;
};

__device__ inline const Slot& part_of(const StradArgs& a, uint32_t s, uint32_t tf, uint32_t u) {
    return u == tf ? a.slots[s] : a.slots[2 * (size_t)u];
}

__device__ inline void part_dev(const Slot& p, double mu, double thr, double& Q, double& D) {
    double d = p.S / (double)p.n - mu;
    if (fabs(d) < thr) d = 0.0;
    const double nd = (double)p.n * d;
    Q += p.Q + 2.0 * d * p.D + nd * d;
    D += p.D + nd;
}

// Per-hcube results of a cut hcube (pyx:2177-2186); c[0..2] = its
// contributions to the mean, the variance and the sum of sigf.
__device__ inline void strad_close(const IterGeom& g, const StradArgs& a, int32_t h, double n, double S, double Q,
                                   double D, DD& m, DD& v, DD& ss, double* c) {
    const double varc = (n * Q - D * D) / (n - 1.);
    c[0] = S + D;
    c[1] = varc;
    c[2] = 0.0;
    dd_add(m, c[0]);
    dd_add(v, varc);
    if (g.adaptive_strat) {
        if (g.fast_pow) {
            const double sv = pow_fast(fabs(varc), g.beta_half);
            a.sigf[h] = sv;
            c[2] = sv;
            dd_add(ss, sv);
        } else {
            a.sigf[h] = fabs(varc);
        }
    }
    if (g.train_errors) a.errv[h] = fabs(varc);
}

// Workgroup `blk` of the straddler pass (kStradThreads lanes); strad_kernel
// runs it as blockIdx.x, a combined launch may place it at another index.
__device__ __forceinline__ void strad_block(const IterGeom& g, const StradArgs& a, uint32_t blk) {
    __shared__ double acc_l[(kStradThreads / 64) * 6];
    const Scalars* sc = a.sc;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    DD m{0, 0}, v{0, 0}, ss{0, 0};
    if (!(sc->status || sc->overflow)) {
        const uint32_t ntiles = sc->b_ntiles, p0 = sc->b_p0;
        const uint32_t s = blk * kStradThreads + threadIdx.x;
        double c[3] = {0.0, 0.0, 0.0};
        int32_t h = -1;
        uint32_t tf = 0, tl = 0, nh = 0;
        bool own = false;
        if (s < 2 * ntiles) {
            h = a.slots[s].h;
            if (h >= 0) {
                const uint32_t lo = a.off[h], hi = a.off[h + 1];
                tf = s >> 1;
                own = lo >= p0 + tf * kT;  // the hcube starts in this tile
                tl = (hi - 1 - p0) / kT;
                nh = hi - lo;
                // a part list must end inside the batch's tiles
                if (own && (hi <= lo || tl >= ntiles)) {
                    atomicOr(&a.sc->status, (int32_t)VGX_ERANGE);
                    own = false;
                }
            }
        }
        const bool in_lane = own && tl - tf < (uint32_t)kLaneParts;
        if (in_lane) {
            double S = 0.0;
            for (uint32_t u = tf; u <= tl; ++u) S += part_of(a, s, tf, u).S;
            const double n = (double)nh, mu = S / n, thr = kEpsilon * fabs(mu);
            double Q = 0.0, D = 0.0;
            for (uint32_t u = tf; u <= tl; ++u) part_dev(part_of(a, s, tf, u), mu, thr, Q, D);
            strad_close(g, a, h, n, S, Q, D, m, v, ss, c);
        }
        // owners of many parts: one at a time across the wave, fixed order
        uint64_t wide = __ballot(own && !in_lane);
        while (wide) {
            const int src = __ffsll((unsigned long long)wide) - 1;
            wide &= wide - 1;
            const int32_t hh = __shfl(h, src, kWave);
            const uint32_t ss_ = __shfl(s, src, kWave), tf_ = __shfl(tf, src, kWave);
            const uint32_t tl_ = __shfl(tl, src, kWave), nh_ = __shfl(nh, src, kWave);
            double S = 0.0;
            for (uint32_t u = tf_ + lane; u <= tl_; u += kWave) S += part_of(a, ss_, tf_, u).S;
            S = wave_sum_fixed(S);
            const double n = (double)nh_, mu = S / n, thr = kEpsilon * fabs(mu);
            double Q = 0.0, D = 0.0;
            for (uint32_t u = tf_ + lane; u <= tl_; u += kWave) part_dev(part_of(a, ss_, tf_, u), mu, thr, Q, D);
            Q = wave_sum_fixed(Q);
            D = wave_sum_fixed(D);
            if (lane == src) strad_close(g, a, hh, n, S, Q, D, m, v, ss, c);
        }
        if (a.cont && s < 2 * ntiles) {
            a.cont[3 * (size_t)s] = c[0];
            a.cont[3 * (size_t)s + 1] = c[1];
            a.cont[3 * (size_t)s + 2] = c[2];
        }
    }
    m = wave_sum_dd(m);
    v = wave_sum_dd(v);
    ss = wave_sum_dd(ss);
    if (lane == 0) {
        acc_l[6 * w + 0] = m.hi; acc_l[6 * w + 1] = m.lo;
        acc_l[6 * w + 2] = v.hi; acc_l[6 * w + 3] = v.lo;
        acc_l[6 * w + 4] = ss.hi; acc_l[6 * w + 5] = ss.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD M{0, 0}, V{0, 0}, T{0, 0};
        for (int k = 0; k < kStradThreads / 64; ++k) {
            M = dd_sum(M, DD{acc_l[6 * k], acc_l[6 * k + 1]});
            V = dd_sum(V, DD{acc_l[6 * k + 2], acc_l[6 * k + 3]});
            T = dd_sum(T, DD{acc_l[6 * k + 4], acc_l[6 * k + 5]});
        }
        double* o = a.strad_part + 6 * (size_t)blk;
        o[0] = M.hi; o[1] = M.lo; o[2] = V.hi; o[3] = V.lo; o[4] = T.hi; o[5] = T.lo;
    }
}

static __global__ __launch_bounds__(kStradThreads) void strad_kernel(IterGeom g, StradArgs a) {
    strad_block(g, a, blockIdx.x);
}

}  // namespace vgx
