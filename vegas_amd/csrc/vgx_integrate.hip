// vgx_integrate.hip -- the vegas+ iteration driver (Integrator.__call__,
// src/vegas/_vegas.pyx:2086-2227) and its non-fused kernels.
//
// Per iteration, all on one stream, no host synchronisation:
//   alloc, offsets        n_h by the damped-variance rule (pyx:1657-1706),
//                         exclusive scan to point offsets, tile -> hcube
//   fused                 sample, map, integrand, per-hcube sums and
//                         sigf = |var|^(beta/2) (pyx:2183-2186), histogram
//   strad                 the same per-hcube sums for hcubes cut by tile edges
//   sigf_pow              exact mode only: sigf by FP64 pow, and its sum
//   finalize              iteration sums (compensated, fixed order), estimate,
//                         sum_sigf (pyx:2209-2215), running average
//                         (pyx:2392-2410); resets the per-iteration scalars
//   adapt                 AdaptiveMap.adapt(alpha) in place (pyx:2216-2217);
//                         clears the training data for the next iteration
// The run starts with iter_init and ends with sigf_reset.
#include <math.h>
#include <string.h>

#include "vgx_fused.h"
#include "vgx_strad.h"
#include "vgx_adapt.h"

namespace vgx {

// Resets the per-iteration scalars (the NaN flag too on a new run) and
// clears the training data (AdaptiveMap.clear, pyx:596-599).
__global__ void iter_init_kernel(IterGeom g, double* sum_f, double* n_f, Scalars* sc, int new_run) {
    const int tot = g.dim * g.nb;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
        sum_f[i] = 0.0;
        n_f[i] = kTiny;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (new_run) {
            sc->status = 0;  // the NaN flag is sticky within one run only
            sc->reset_sigf = 0;
        }
        sc->wfn_max_bits = 0;
        for (int k = 0; k < 6; ++k) sc->acc[k] = 0.0;
        sc->overflow = 0;
    }
}

// ------------------------------------------------------------ allocation ----
// n_h = min(int(sigf_h * neval_sigf) + nmin, max(nmax, nmin))   (pyx:1696-1698)
// Per-workgroup total, min and max go to stat[3 * blockIdx.x + {0,1,2}].
// The iteration's training data start empty (AdaptiveMap.clear, pyx:596-599):
// they are cleared here rather than after the adapt reads them, so the last
// iteration's histogram stays readable after a run.
__global__ __launch_bounds__(kAllocThreads) void alloc_kernel(IterGeom g, double* sigf,
                                                              const double* sum_sigf, int32_t* nh,
                                                              uint64_t* stat, const Scalars* sc,
                                                              double* sum_f, double* n_f) {
    __shared__ uint64_t red[kAllocThreads / 64];
    __shared__ int32_t rmin[kAllocThreads / 64], rmax[kAllocThreads / 64];
    if (sc->status) return;
    if (g.train_points || g.train_errors) {
        const int tot = g.dim * g.nb;
        for (int i = blockIdx.x * kAllocThreads + threadIdx.x; i < tot; i += gridDim.x * kAllocThreads) {
            sum_f[i] = 0.0;
            n_f[i] = kTiny;
        }
    }
    const int reset = sc->reset_sigf;  // previous iteration found sum(sigf) == 0
    const double ss = *sum_sigf;
    const double neval_sigf = (g.beta > 0 && ss > 0 && g.adaptive_sampler) ? g.neval_frac * g.neval / ss : 0.0;
    uint64_t sum = 0;
    int32_t mn = g.nmin, mx = g.nmin;
    const int64_t base = (int64_t)blockIdx.x * kAllocPerBlock;
    for (int i = 0; i < kAllocItems; ++i) {
        const int64_t h = base + (int64_t)i * kAllocThreads + threadIdx.x;
        if (h >= g.nhcube) break;
        int32_t n;
        if (g.adaptive_sampler) {
            double sv;
            if (reset) {
                sv = 1.0;
                sigf[h] = 1.0;
            } else {
                sv = sigf[h];
            }
            double v = sv * neval_sigf;
            v = v < 2147483647.0 ? v : 2147483647.0;
            int64_t nn = (int64_t)(int32_t)v + g.nmin;
            if (nn > g.nmax) nn = g.nmax;
            n = (int32_t)nn;
            mn = n < mn ? n : mn;
            mx = n > mx ? n : mx;
        } else {
            n = g.avg_nh;
        }
        nh[h] = n;
        sum += (uint64_t)n;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o, kWave);
        mn = min(mn, __shfl_xor(mn, o, kWave));
        mx = max(mx, __shfl_xor(mx, o, kWave));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w] = sum;
        rmin[w] = mn;
        rmax[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int k = 0; k < kAllocThreads / 64; ++k) {
            s += red[k];
            mn = min(mn, rmin[k]);
            mx = max(mx, rmax[k]);
        }
        stat[3 * blockIdx.x] = s;
        stat[3 * blockIdx.x + 1] = (uint64_t)(uint32_t)mn;
        stat[3 * blockIdx.x + 2] = (uint64_t)(uint32_t)mx;
    }
}

// Point offsets off[h]: exclusive scan of n_h.  Each workgroup takes its
// base from the totals of the workgroups before it (alloc_kernel's stat),
// scans its 2048 hcubes through LDS and writes them coalesced, with
// tile -> first hcube.  The last workgroup also writes the iteration's point
// count, neval_hcube_range and the tile count (or flags an overflow of the
// sample buffer).  LDS items are padded (item j at j + j/8) so the per-thread
// runs of 8 are read without bank conflicts.
__device__ inline int lds_pad(int j) { return j + (j >> 3); }

__global__ __launch_bounds__(kAllocThreads) void offsets_kernel(IterGeom g, const int32_t* nh,
                                                                const uint64_t* stat, int nblk, uint32_t* off,
                                                                int32_t* tile_h, Scalars* sc,
                                                                uint64_t capacity) {
    __shared__ uint32_t ln[kAllocPerBlock + kAllocPerBlock / 8];
    __shared__ uint64_t wred[2 * (kAllocThreads / 64)];
    __shared__ int32_t wmm[2 * (kAllocThreads / 64)];
    __shared__ uint32_t wsum[kAllocThreads / 64];
    if (sc->status) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int nw = kAllocThreads / 64;
    const int64_t base = (int64_t)blockIdx.x * kAllocPerBlock;
    // this workgroup's n_h into LDS first: those loads are in flight while
    // the totals below are read and reduced (the barrier after the wave
    // reductions orders them)
    for (int i = threadIdx.x; i < kAllocPerBlock; i += kAllocThreads) {
        const int64_t h = base + i;
        ln[lds_pad(i)] = h < g.nhcube ? (uint32_t)nh[h] : 0u;
    }
    // totals before this workgroup and over all workgroups
    uint64_t pre = 0, tot = 0;
    int32_t mn = g.nmin, mx = g.nmin;
    const bool last = blockIdx.x == gridDim.x - 1;
    for (int i = threadIdx.x; i < nblk; i += kAllocThreads) {
        const uint64_t v = stat[3 * i];
        tot += v;
        if (i < (int)blockIdx.x) pre += v;
        if (last) {
            mn = min(mn, (int32_t)stat[3 * i + 1]);
            mx = max(mx, (int32_t)stat[3 * i + 2]);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pre += __shfl_xor(pre, o, kWave);
        tot += __shfl_xor(tot, o, kWave);
        mn = min(mn, __shfl_xor(mn, o, kWave));
        mx = max(mx, __shfl_xor(mx, o, kWave));
    }
    if (lane == 0) {
        wred[2 * w] = pre;
        wred[2 * w + 1] = tot;
        wmm[2 * w] = mn;
        wmm[2 * w + 1] = mx;
    }
    __syncthreads();
    pre = 0;
    tot = 0;
    for (int k = 0; k < nw; ++k) {
        pre += wred[2 * k];
        tot += wred[2 * k + 1];
        mn = min(mn, wmm[2 * k]);
        mx = max(mx, wmm[2 * k + 1]);
    }
    const bool ok = tot <= capacity && tot < 0xFFFFFFFFull;
    if (last && threadIdx.x == 0) {
        sc->overflow = ok ? 0u : 1u;
        sc->total = ok ? (uint32_t)tot : 0u;
        sc->nh_min = mn;
        sc->nh_max = mx;
        sc->b_p0 = 0;
        sc->b_p1 = sc->total;
        sc->b_h0 = 0;
        sc->b_h1 = (uint32_t)g.nhcube;
        sc->b_ntiles = (sc->total + kT - 1) / kT;
    }
    if (!ok) return;
    uint32_t loc[kAllocItems];
    uint32_t run = 0;
    for (int i = 0; i < kAllocItems; ++i) {
        loc[i] = run;
        run += ln[lds_pad(threadIdx.x * kAllocItems + i)];
    }
    uint32_t incl = run;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t wbase = 0;
    for (int k = 0; k < w; ++k) wbase += wsum[k];
    const uint32_t tbase = (uint32_t)pre + wbase + incl - run;
    for (int i = 0; i < kAllocItems; ++i) {
        const int64_t h = base + threadIdx.x * kAllocItems + i;
        const uint32_t lo = tbase + loc[i];
        const uint32_t n = ln[lds_pad(threadIdx.x * kAllocItems + i)];
        if (h < g.nhcube) {
            const uint32_t hi = lo + n;
            for (uint32_t t = (lo + kT - 1) / kT; t < (hi + kT - 1) / kT; ++t) tile_h[t] = (int32_t)h;
        }
        loc[i] = lo;
    }
    __syncthreads();
    for (int i = 0; i < kAllocItems; ++i) ln[lds_pad(threadIdx.x * kAllocItems + i)] = loc[i];
    __syncthreads();
    for (int i = threadIdx.x; i < kAllocPerBlock; i += kAllocThreads) {
        const int64_t h = base + i;
        if (h < g.nhcube) off[h] = ln[lds_pad(i)];
    }
    if (last && threadIdx.x == 0) off[g.nhcube] = (uint32_t)tot;
}

// Staged batches: tiles of points [p0, p1) covering hcubes [h0, h1).
__global__ void batch_tiles_kernel(const uint32_t* off, int32_t* tile_h, Scalars* sc, uint32_t h0,
                                   uint32_t h1, uint32_t p0, uint32_t p1) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->b_p0 = p0;
        sc->b_p1 = p1;
        sc->b_h0 = h0;
        sc->b_h1 = h1;
        sc->b_ntiles = (p1 - p0 + kT - 1) / kT;
    }
    for (uint32_t h = h0 + blockIdx.x * blockDim.x + threadIdx.x; h < h1; h += gridDim.x * blockDim.x) {
        const uint32_t lo = off[h] - p0, hi = off[h + 1] - p0;
        for (uint32_t t = (lo + kT - 1) / kT; t < (hi + kT - 1) / kT; ++t) tile_h[t] = (int32_t)h;
    }
}

constexpr int kReduceThreads = 256;  // threads of the reduction / finalize code

// Fixed-order tree over DD triples: red[k*6 .. k*6+5] for k < blockDim.x.
__device__ void tree_dd3(double* red) {
    for (int o = kReduceThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) {
            const int i = threadIdx.x, j = threadIdx.x + o;
            for (int q = 0; q < 3; ++q) {
                DD r = dd_sum(DD{red[6 * i + 2 * q], red[6 * i + 2 * q + 1]},
                              DD{red[6 * j + 2 * q], red[6 * j + 2 * q + 1]});
                red[6 * i + 2 * q] = r.hi;
                red[6 * i + 2 * q + 1] = r.lo;
            }
        }
        __syncthreads();
    }
}

// Adds partial sums (fused workgroups, then straddler workgroups, fixed
// order) into the iteration accumulators sc->acc (DD mean, var, sigf sum).
// One workgroup of kReduceThreads.
__device__ void reduce_parts(const double* blk_part, int nblk, const double* strad_part, int nstrad,
                             Scalars* sc, double* red) {
    DD m{0, 0}, v{0, 0}, t{0, 0};
    const int tot = nblk + nstrad;
    for (int i = threadIdx.x; i < tot; i += kReduceThreads) {
        const double* p = i < nblk ? blk_part + 6 * (size_t)i : strad_part + 6 * (size_t)(i - nblk);
        m = dd_sum(m, DD{p[0], p[1]});
        v = dd_sum(v, DD{p[2], p[3]});
        t = dd_sum(t, DD{p[4], p[5]});
    }
    red[6 * threadIdx.x + 0] = m.hi; red[6 * threadIdx.x + 1] = m.lo;
    red[6 * threadIdx.x + 2] = v.hi; red[6 * threadIdx.x + 3] = v.lo;
    red[6 * threadIdx.x + 4] = t.hi; red[6 * threadIdx.x + 5] = t.lo;
    __syncthreads();
    tree_dd3(red);
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q) {
            const DD r = dd_sum(DD{sc->acc[2 * q], sc->acc[2 * q + 1]}, DD{red[2 * q], red[2 * q + 1]});
            sc->acc[2 * q] = r.hi;
            sc->acc[2 * q + 1] = r.lo;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kReduceThreads) void batch_reduce_kernel(const double* blk_part, int nblk,
                                                                      const double* strad_part, int nstrad,
                                                                      Scalars* sc) {
    __shared__ double red[kReduceThreads * 6];
    if (sc->status || sc->overflow) return;
    reduce_parts(blk_part, nblk, strad_part, nstrad, sc, red);
}

// sigf_h = |var_h|^(beta/2) (pyx:2185) and per-workgroup sums, after every
// hcube of the iteration is done.
__global__ __launch_bounds__(256) void sigf_pow_kernel(IterGeom g, double* sigf, double* part, const Scalars* sc) {
    __shared__ double red[8];
    if (sc->status || sc->overflow) return;
    DD s{0, 0};
    for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < g.nhcube;
         h += (int64_t)gridDim.x * blockDim.x) {
        // exact mode: pow as the reference; otherwise exp2(b log2 x), within
        // ~1e-14 relative (sigf only steers the allocation)
        const double x = sigf[h];
        const double v = g.exact_div ? pow(x, g.beta_half) : (x > 0 ? exp2(g.beta_half * log2(x)) : 0.0);
        sigf[h] = v;
        dd_add(s, v);
    }
    s = wave_sum_dd(s);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[2 * w] = s.hi;
        red[2 * w + 1] = s.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t{0, 0};
        for (int k = 0; k < 4; ++k) t = dd_sum(t, DD{red[2 * k], red[2 * k + 1]});
        part[2 * blockIdx.x] = t.hi;
        part[2 * blockIdx.x + 1] = t.lo;
    }
}

// Iteration close (one workgroup): the batch's partial sums when given
// (single-batch iterations), the sum of sigf (accumulated with the sums, or
// from the sigf_pow partials in exact mode), the estimate, sum_sigf bookkeeping and the FP64 running average; resets the
// per-iteration scalars for the next iteration of the run.
// Runs on threads 0..kReduceThreads-1 of the calling workgroup (the others
// must have returned); red: kReduceThreads * 6 doubles of LDS, mv: 2.
__device__ void finalize_body(const IterGeom& g, Scalars* sc, const vgx_state& st, int has_sigf,
                              const double* sigf_part, int nsigf, const double* blk_part, int nblk,
                              const double* strad_part, int nstrad, double* red, double* mv) {
    if (nblk + nstrad > 0 && !(sc->status || sc->overflow))
        reduce_parts(blk_part, nblk, strad_part, nstrad, sc, red);
    const bool sig_parts = g.adaptive_strat && !g.fast_pow;
    if (sig_parts) {
        DD t{0, 0};
        for (int k = threadIdx.x; k < nsigf; k += kReduceThreads)
            t = dd_sum(t, DD{sigf_part[2 * k], sigf_part[2 * k + 1]});
        red[2 * threadIdx.x] = t.hi;
        red[2 * threadIdx.x + 1] = t.lo;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < (unsigned)o) {
                DD r = dd_sum(DD{red[2 * threadIdx.x], red[2 * threadIdx.x + 1]},
                              DD{red[2 * (threadIdx.x + o)], red[2 * (threadIdx.x + o) + 1]});
                red[2 * threadIdx.x] = r.hi;
                red[2 * threadIdx.x + 1] = r.lo;
            }
            __syncthreads();
        }
    }
    const double ssig_parts = sig_parts ? red[0] + red[1] : 0.0;
    __syncthreads();
    // The workgroup stops here on an error; every thread reads the same flags.
    const int ovf = sc->overflow, bad = sc->status;
    if (ovf || bad) {
        if (threadIdx.x == 0) {
            if (st.status) *st.status = ovf ? VGX_ERANGE : bad;
            if (ovf) sc->status = VGX_ERANGE;
        }
        return;
    }
    const int it = (int)g.itn;
    if (threadIdx.x == 0) {
        const double mean = sc->acc[0] + sc->acc[1];
        const double var = sc->acc[2] + sc->acc[3];
        mv[0] = mean;
        mv[1] = var;
        const double ssig = g.fast_pow ? sc->acc[4] + sc->acc[5] : ssig_parts;
        st.itn_mean[it] = mean;
        st.itn_var[it] = var;
        st.itn_neval[it] = (int64_t)sc->total;
        st.itn_nhrange[2 * it] = g.adaptive_sampler ? sc->nh_min : g.nmin;
        st.itn_nhrange[2 * it + 1] = g.adaptive_sampler ? sc->nh_max : g.nmin;
        sc->reset_sigf = 0;
        // sum_sigf (pyx:2209-2215); sigf[:] = 1 is applied lazily by the next
        // allocation, and by sigf_reset_kernel at the end of the run
        if (g.beta > 0 && g.adapt && !g.train_errors) {
            if (ssig > 0) {
                *st.sum_sigf = ssig;
            } else {
                sc->reset_sigf = has_sigf;
                *st.sum_sigf = has_sigf ? (double)g.nhcube : 0.0;
            }
        }
    }
    // running average over iterations 0..it (RAvg.add, pyx:2392-2410): the
    // earlier iterations' (mean, var) are staged through LDS in chunks (one
    // load latency per chunk) and summed by thread 0 in iteration order
    const int n = it + 1;
    constexpr int kChunk = kReduceThreads * 3;  // (mean, var) pairs in red[]
    auto stage = [&](int k0) {
        __syncthreads();
        for (int k = k0 + (int)threadIdx.x; k < min(n, k0 + kChunk); k += kReduceThreads) {
            red[2 * (k - k0)] = k == it ? mv[0] : st.itn_mean[k];
            red[2 * (k - k0) + 1] = k == it ? mv[1] : st.itn_var[k];
        }
        __syncthreads();
    };
    double rm = 0.0, rs = 0.0, chi2 = 0.0;
    if (g.weighted) {
        double sw = 0.0, swm = 0.0;
        for (int k0 = 0; k0 < n; k0 += kChunk) {
            stage(k0);
            if (threadIdx.x == 0)
                for (int k = k0; k < min(n, k0 + kChunk); ++k) {
                    const double sd = sqrt(red[2 * (k - k0) + 1]);
                    const double vk = sd * sd;
                    const double w = 1. / (vk > kTiny ? vk : kTiny);
                    sw += w;
                    swm += w * red[2 * (k - k0)];
                }
        }
        const double v = 1. / sw;
        rm = swm * v;
        rs = sqrt(v);
        if (n > 1)
            for (int k0 = 0; k0 < n; k0 += kChunk) {
                if (n > kChunk) stage(k0);
                if (threadIdx.x == 0)
                    for (int k = k0; k < min(n, k0 + kChunk); ++k) {
                        const double sd = sqrt(red[2 * (k - k0) + 1]);
                        const double vk = sd * sd;
                        const double w = 1. / (vk > kTiny ? vk : kTiny);
                        const double dm = rm - red[2 * (k - k0)];
                        chi2 += dm * dm * w;
                    }
            }
    } else {
        double ms = 0.0, vs = 0.0;
        for (int k0 = 0; k0 < n; k0 += kChunk) {
            stage(k0);
            if (threadIdx.x == 0)
                for (int k = k0; k < min(n, k0 + kChunk); ++k) {
                    const double sd = sqrt(red[2 * (k - k0) + 1]);
                    ms += red[2 * (k - k0)];
                    vs += sd * sd;
                }
        }
        rm = ms / n;
        rs = sqrt(vs / ((double)n * (double)n));
        if (n > 1) {
            for (int k0 = 0; k0 < n; k0 += kChunk) {
                if (n > kChunk) stage(k0);
                if (threadIdx.x == 0)
                    for (int k = k0; k < min(n, k0 + kChunk); ++k) {
                        const double dm = red[2 * (k - k0)] - rm;
                        chi2 += dm * dm;
                    }
            }
            chi2 /= (vs / n);
        }
    }
    if (threadIdx.x != 0) return;
    st.ravg[0] = rm;
    st.ravg[1] = rs;
    st.ravg[2] = chi2;
    st.ravg[3] = (double)(n - 1);
    // next iteration of this run
    sc->wfn_max_bits = 0;
    for (int k = 0; k < 6; ++k) sc->acc[k] = 0.0;
}

__global__ __launch_bounds__(kReduceThreads) void finalize_kernel(IterGeom g, Scalars* sc, vgx_state st,
                                                                  int has_sigf, const double* sigf_part,
                                                                  int nsigf, const double* blk_part, int nblk,
                                                                  const double* strad_part, int nstrad) {
    __shared__ double red[kReduceThreads * 6];
    __shared__ double mv[2];
    finalize_body(g, sc, st, has_sigf, sigf_part, nsigf, blk_part, nblk, strad_part, nstrad, red, mv);
}

// Iteration close with the parallel adapt in one launch: workgroups
// 0..dim-1 adapt one axis each (they read only the training histogram and
// the grid), workgroup dim runs finalize_body on its first kReduceThreads
// threads (it touches neither).  Waves that return early leave the
// workgroup's barriers to the others.
__global__ __launch_bounds__(1024) void close_kernel(IterGeom g, Scalars* sc, vgx_state st, int has_sigf,
                                                     const double* sigf_part, int nsigf, const double* blk_part,
                                                     int nblk, const double* strad_part, int nstrad,
                                                     AdaptArgs a) {
    if ((int)blockIdx.x == a.dim) {
        __shared__ double red[kReduceThreads * 6];
        __shared__ double mv[2];
        if (threadIdx.x >= (unsigned)kReduceThreads) return;
        finalize_body(g, sc, st, has_sigf, sigf_part, nsigf, blk_part, nblk, strad_part, nstrad, red, mv);
        return;
    }
    __shared__ double av[kAdaptLdsMax];
    __shared__ double P[kAdaptLdsMax];
    __shared__ double ared[40];
    // the flags finalize_body may raise are set before this launch
    if (sc->status || sc->overflow) return;
    adapt_parallel_axis(a, blockIdx.x, av, P, ared);
}

// The training histogram of one fused launch: its workgroups' rows summed
// in row order (row group j takes rows j, j + kHrGroups, ...; the groups are
// added in order) and added to sum_f / n_f.  Each thread issues its loads in
// batches of kHrBatch rows so the device keeps many in flight.  Runs as
// extra workgroups of the straddler launch (the two are independent).
constexpr int kHrCols = 32, kHrGroups = kStradThreads / kHrCols, kHrBatch = 8;

struct HistArgs {
    const double* ps;     // [nrows][tot] sums
    const uint32_t* pc;   // [nrows][tot] counts
    int nrows, tot;
    double* sum_f;
    double* n_f;
};

__device__ __forceinline__ void hist_reduce_block(const HistArgs& h, const Scalars* sc, uint32_t blk) {
    __shared__ double rs[kHrGroups][kHrCols];
    __shared__ uint64_t rn[kHrGroups][kHrCols];
    __shared__ int skip;
    // the sampler did not run, or its rows are void; read once per workgroup
    // (straddler workgroups of the same launch may raise the status meanwhile)
    if (threadIdx.x == 0) skip = (sc->status || sc->overflow) ? 1 : 0;
    __syncthreads();
    if (skip) return;
    const int il = threadIdx.x % kHrCols, j = threadIdx.x / kHrCols;
    const int i = (int)blk * kHrCols + il;
    double s = 0.0;
    uint64_t n = 0;
    if (i < h.tot) {
        int w = j;
        for (; w + (kHrBatch - 1) * kHrGroups < h.nrows; w += kHrBatch * kHrGroups) {
            double v[kHrBatch];
            uint32_t c[kHrBatch];
#pragma unroll
            for (int b = 0; b < kHrBatch; ++b) {
                v[b] = h.ps[(size_t)(w + b * kHrGroups) * h.tot + i];
                c[b] = h.pc[(size_t)(w + b * kHrGroups) * h.tot + i];
            }
#pragma unroll
            for (int b = 0; b < kHrBatch; ++b) {
                s += v[b];
                n += c[b];
            }
        }
        for (; w < h.nrows; w += kHrGroups) {
            s += h.ps[(size_t)w * h.tot + i];
            n += h.pc[(size_t)w * h.tot + i];
        }
    }
    rs[j][il] = s;
    rn[j][il] = n;
    __syncthreads();
    if (j == 0 && i < h.tot) {
        double t = 0.0;
        uint64_t m = 0;
        for (int q = 0; q < kHrGroups; ++q) {
            t += rs[q][il];
            m += rn[q][il];
        }
        h.sum_f[i] += t;
        h.n_f[i] += (double)m;
    }
}

// Training histogram of a fast-mode batch from the fused sampler's records,
// in a fixed order.  The batch's points are cut into kTrainBlocks contiguous
// blocks (a multiple of 64 points each); workgroup (block b, axis d) is one
// wave that adds its block's points into an LDS row of axis d (FP64 sums,
// u32 counts) and writes that row to part d of row b, which
// hist_reduce_block sums over the blocks in a fixed order.  Each bin is thus
// only ever added to by one wave, in a fixed order (rounds of 256 points,
// then the 4 points of a lane, then the lanes as the LDS orders one
// instruction's lanes), so a run's histogram does not depend on timing.
// A lane reads 4 consecutive points per round (one 16-byte load of their
// bin words, two of their values), in batches of kTrainRounds rounds whose
// loads are issued a batch ahead of their additions.  Workgroups are dealt to the 8 XCDs round-robin by id, so
// the id keeps the XCD slot of the block (id mod 8 = b mod 8): the waves of
// one block run on one XCD and share the records' lines in its L2.
constexpr int kTrainBlocks = 256, kTrainRounds = 8, kXcds = 8;
static_assert(kTrainBlocks % kXcds == 0, "blocks must fill the XCD slots evenly");

// kTrainRounds rounds of 256 points of one axis: a lane's 4 bins and values
struct TrainBatch {
    uint4 w[kTrainRounds];
    double2 c01[kTrainRounds], c23[kTrainRounds];
};

__device__ __forceinline__ void train_load(TrainBatch& t, const uint32_t* words, const double* rc2, uint32_t p,
                                           uint32_t hi, int lane) {
#pragma unroll
    for (int r = 0; r < kTrainRounds; ++r) {
        const uint32_t i = p + 256u * r + 4u * lane;  // a multiple of 4 (p is)
        if (i + 3 < hi) {
            t.w[r] = *reinterpret_cast<const uint4*>(words + i);
            t.c01[r] = *reinterpret_cast<const double2*>(rc2 + i);
            t.c23[r] = *reinterpret_cast<const double2*>(rc2 + i + 2);
        } else {
            const uint32_t none = kNoBin | (kNoBin << 16);
            t.w[r] = uint4{i < hi ? words[i] : none, i + 1 < hi ? words[i + 1] : none,
                           i + 2 < hi ? words[i + 2] : none, none};
            t.c01[r] = double2{i < hi ? rc2[i] : 0.0, i + 1 < hi ? rc2[i + 1] : 0.0};
            t.c23[r] = double2{i + 2 < hi ? rc2[i + 2] : 0.0, 0.0};
        }
    }
}

__device__ __forceinline__ void train_add(const TrainBatch& t, int sh, uint32_t ni, double* Ls, uint32_t* Lc) {
#pragma unroll
    for (int r = 0; r < kTrainRounds; ++r) {
        const uint32_t k4[4] = {(t.w[r].x >> sh) & 0xFFFFu, (t.w[r].y >> sh) & 0xFFFFu,
                                (t.w[r].z >> sh) & 0xFFFFu, (t.w[r].w >> sh) & 0xFFFFu};
        const double v4[4] = {t.c01[r].x, t.c01[r].y, t.c23[r].x, t.c23[r].y};
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k4[j] < ni) {  // kNoBin and any bin past ninc are skipped
                unsafeAtomicAdd(&Ls[k4[j]], v4[j]);
                atomicAdd(&Lc[k4[j]], 1u);
            }
    }
}

__global__ __launch_bounds__(64) void train_rows_kernel(IterGeom g, const uint32_t* rbins, const double* rc2,
                                                        uint32_t rstride, double* rows_s, uint32_t* rows_c,
                                                        const Scalars* sc) {
    extern __shared__ double trow[];
    const int dim = g.dim, nb = g.nb;
    const int xcd = (int)blockIdx.x % kXcds, rest = (int)blockIdx.x / kXcds;
    const int d = rest % dim, b = (rest / dim) * kXcds + xcd;
    const int lane = threadIdx.x;
    double* Ls = trow;
    uint32_t* Lc = reinterpret_cast<uint32_t*>(trow + nb);
    for (int k = lane; k < nb; k += 64) {
        Ls[k] = 0.0;
        Lc[k] = 0u;
    }
    __syncthreads();
    // the sampler did not run, or its records are void: the rows stay zero
    const bool skip = sc->status || sc->overflow;
    const uint32_t ni = (uint32_t)g.ninc[d];
    const uint32_t* words = rbins + (size_t)(d >> 1) * rstride;
    const int sh = 16 * (d & 1);
    if (!skip) {
        const uint32_t n = sc->b_p1 - sc->b_p0;
        const uint32_t chunk = ((n + kTrainBlocks - 1) / kTrainBlocks + 63u) & ~63u;
        const uint32_t lo = min((uint32_t)b * chunk, n), hi = min(lo + chunk, n);
        constexpr uint32_t step = 256u * kTrainRounds;
        // software pipelined: the loads of the next batch are in flight
        // while the current one is added (the additions keep their order)
        TrainBatch A, B;
        if (lo < hi) train_load(A, words, rc2, lo, hi, lane);
        for (uint32_t p = lo; p < hi; p += 2 * step) {
            if (p + step < hi) train_load(B, words, rc2, p + step, hi, lane);
            train_add(A, sh, ni, Ls, Lc);
            if (p + step >= hi) break;
            if (p + 2 * step < hi) train_load(A, words, rc2, p + 2 * step, hi, lane);
            train_add(B, sh, ni, Ls, Lc);
        }
    }
    __syncthreads();
    const size_t tot = (size_t)dim * nb;
    double* rs = rows_s + (size_t)b * tot + (size_t)d * nb;
    uint32_t* rcn = rows_c + (size_t)b * tot + (size_t)d * nb;
    for (int k = lane; k < nb; k += 64) {
        rs[k] = Ls[k];
        rcn[k] = Lc[k];
    }
}

// Straddler workgroups 0..nstrad-1, histogram-row workgroups after them.
__global__ __launch_bounds__(kStradThreads) void strad_hist_kernel(IterGeom g, StradArgs a, HistArgs h,
                                                                    int nstrad) {
    if ((int)blockIdx.x < nstrad)
        strad_block(g, a, blockIdx.x);
    else
        hist_reduce_block(h, a.sc, blockIdx.x - (uint32_t)nstrad);
}

__global__ void sigf_reset_kernel(double* sigf, int64_t n, Scalars* sc) {
    if (!sc->reset_sigf || sc->status) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        sigf[i] = 1.0;
}

// eval_point for hcube h with the stratum->bin path the fused sampler takes
// for this geometry (staged and error-training kernels).
template <bool R32>
__device__ inline void eval_point_h(const IterGeom& g, const AxisC* axes, const double* grid, uint32_t h,
                                    uint32_t p, double* x, double& J, int* kb) {
    PackedCoords<0> pc;
    uint32_t zmin = 0xFFFFFFFFu;
    if constexpr (R32) {
        if (g.per_all) {
            pc.set(axes, g.dim, h);
            eval_point<0, true, true>(tile_geom(g), axes, grid, pc, h, p, x, J, kb, zmin);
            if (zmin == 0) {
                const uint64_t zm = per_zero_axes<0>(tile_geom(g), pc, p);
                for (int d = 0; d < g.dim; ++d)
                    if ((zm >> d) & 1) kb[d] = -1;
            }
            return;
        }
    }
    eval_point<0, R32, false>(tile_geom(g), axes, grid, pc, h, p, x, J, kb, zmin);
}

// adapt_to_errors: training entry |var_h| at the last point of every hcube
// (pyx:2187-2193).
template <bool R32>
__global__ __launch_bounds__(256) void errors_train_kernel(IterGeom g, const AxisC* axes, const double* grid,
                                                           const uint32_t* off, const double* errv,
                                                           double* sum_f, double* n_f, const Scalars* sc) {
    if (sc->status || sc->overflow) return;
    for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < g.nhcube;
         h += (int64_t)gridDim.x * blockDim.x) {
        double x[VGX_MAXDIM];
        int kb[VGX_MAXDIM];
        double J;
        eval_point_h<R32>(g, axes, grid, (uint32_t)h, off[h + 1] - 1, x, J, kb);
        for (int d = 0; d < g.dim; ++d)
            if (kb[d] >= 0) {
                unsafeAtomicAdd(&sum_f[d * g.nb + kb[d]], errv[h]);
                unsafeAtomicAdd(&n_f[d * g.nb + kb[d]], 1.0);
            }
    }
}

__global__ void axes_init_kernel(IterGeom g, AxisC* axes) {
    if (threadIdx.x < (unsigned)g.dim) axes[threadIdx.x] = make_axis(g, threadIdx.x);
}

__global__ void to_float_kernel(const double* a, float* b, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)a[i];
}

// Staged path: x, y, weight and hcube of points [p0, p1) (pyx:1731-1759).
template <bool R32>
__global__ __launch_bounds__(256) void sample_kernel(IterGeom g, const AxisC* axes, const double* grid,
                                                     const uint32_t* off, uint32_t h0, uint32_t h1,
                                                     uint32_t p0, uint32_t p1, double* x, double* y,
                                                     double* wgt, int64_t* hcube, const Scalars* sc) {
    if (sc->status) return;
    const int dim = g.dim;
    for (uint32_t p = p0 + blockIdx.x * blockDim.x + threadIdx.x; p < p1; p += gridDim.x * blockDim.x) {
        uint32_t lo = h0, hi = h1 - 1;  // largest h with off[h] <= p
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (off[mid] <= p) lo = mid; else hi = mid - 1;
        }
        const uint32_t h = lo;
        const uint32_t n = off[h + 1] - off[h];
        double xx[VGX_MAXDIM];
        int kb[VGX_MAXDIM];
        double J;
        eval_point_h<R32>(g, axes, grid, h, p, xx, J, kb);
        const int64_t i = (int64_t)(p - p0);
        for (int d = 0; d < dim; ++d) x[i * dim + d] = xx[d];
        if (y) point_y<R32>(tile_geom(g), axes, h, p, y + i * dim);
        if (wgt) wgt[i] = J * (g.dv_y / (double)n);
        if (hcube) hcube[i] = (int64_t)h;
    }
}

// ------------------------------------------------------------ host side ----
static int build_geom(const vgx_config* c, const DeviceInfo& info, IterGeom* g) {
    memset(g, 0, sizeof(*g));
    if (c->dim < 1 || c->dim > VGX_MAXDIM || c->nitn < 0) return VGX_EINVAL;
    if (c->nhcube < 1 || c->nhcube >= 0x7FFFFFFFll) return VGX_ERANGE;
    g->dim = c->dim;
    g->ld = c->ld;
    g->nb = c->ld - 1;
    int64_t nh = 1;
    for (int d = 0; d < c->dim; ++d) {
        if (c->ninc[d] < 1 || c->ninc[d] + 1 > c->ld || c->nstrat[d] < 1) return VGX_EINVAL;
        g->ninc[d] = c->ninc[d];
        g->nstrat[d] = c->nstrat[d];
        g->inv_nstrat[d] = 1.0 / (double)c->nstrat[d];
        g->dm[d] = make_divmagic((uint32_t)c->nstrat[d]);
        nh *= c->nstrat[d];
    }
    if (nh != c->nhcube) return VGX_EINVAL;
    g->nhcube = c->nhcube;
    g->neval = (double)c->neval;
    g->neval_frac = c->neval_frac;
    g->beta = c->beta;
    g->beta_half = c->beta / 2.;
    g->dv_y = 1. / (double)c->nhcube;
    g->adaptive_sampler = c->beta > 0 && c->nhcube > 1 && !c->adapt_to_errors;
    g->adaptive_strat = g->adaptive_sampler && c->adapt;
    g->train_points = !c->adapt_to_errors && c->adapt && c->alpha > 0;
    g->train_errors = c->adapt_to_errors && c->adapt;
    g->rng32 = c->precision == VGX_PREC_F32;
    g->exact_div = (c->flags & VGX_FLAG_EXACT_DIV) ? 1 : 0;
    g->weighted = c->weighted;
    g->adapt = c->adapt;
    g->fast_pow = g->adaptive_strat && !g->exact_div;
    g->per_all = g->rng32 && !g->exact_div;
    for (int d = 0; d < c->dim; ++d)
        if (make_axis(*g, d).per == 0 || c->ninc[d] >= 65536) g->per_all = 0;
    g->jn = 1.0;
    for (int d = 0; d < c->dim; ++d) g->jn *= (double)c->ninc[d];
    const int64_t nmin = c->min_neval_hcube;
    const int64_t nmax = c->max_neval_hcube > nmin ? c->max_neval_hcube : nmin;
    const int64_t avg = (int64_t)((double)c->neval / (double)c->nhcube);
    if (nmin < 2 || nmax > 0x7FFFFFFF || (!g->adaptive_sampler && avg < 2)) return VGX_ERANGE;
    g->nmin = (int32_t)nmin;
    g->nmax = (int32_t)nmax;
    g->avg_nh = (int32_t)(avg > 0x7FFFFFFF ? 0x7FFFFFFF : avg);
    g->seed_lo = (uint32_t)(c->seed & 0xffffffffu);
    g->seed_hi = (uint32_t)(c->seed >> 32);
    g->nitn = (uint32_t)c->nitn;
    // the caller's uniforms take the place of the 53-bit Philox stream for one
    // staged iteration
    if (c->uniforms && (c->precision != VGX_PREC_F64 || c->nitn != 1)) return VGX_EINVAL;
    g->uext = c->uniforms;
    // LDS plan of the fused kernel: the FP64 grid, offsets slices, and in
    // fast mode the training histogram (FP64 sums, u16 pair counts) when it
    // fits; otherwise (and in exact mode) the histogram is built from the
    // training records after the launch.  `fixed` includes the kernel's
    // static LDS (flags, the tile counter and the adder state, under 1 KiB).
    const size_t tot = (size_t)c->dim * g->nb;
    const size_t fixed = (size_t)16 * kNoff * sizeof(uint32_t) + 1024;
    const size_t gbytes = (size_t)c->dim * c->ld * sizeof(double);
    const size_t hbytes = tot * sizeof(double) + ((tot + 1) / 2) * sizeof(uint32_t) + 32;
    g->grid_lds = 1;
    if (fixed + gbytes > info.max_lds) return VGX_ERANGE;
    // (the ordered LDS histogram is in the integer stratum path's kernels)
    g->hist_lds = g->train_points && g->per_all && !(c->flags & VGX_FLAG_EXACT_ADAPT) &&
                  (fixed + gbytes + hbytes <= info.max_lds);
    g->hist_unordered = g->hist_lds && (c->flags & VGX_FLAG_UNORDERED_HIST) ? 1 : 0;
    // the training-row kernel's LDS row of one axis
    if ((size_t)g->nb * (sizeof(double) + sizeof(uint32_t)) > info.max_lds) return VGX_ERANGE;
    return VGX_OK;
}

static size_t fused_lds(const IterGeom& g) {
    const size_t tot = (size_t)g.dim * g.nb;
    size_t b = (size_t)g.dim * g.ld * sizeof(double);
    if (g.hist_lds && g.train_points)
        b += tot * sizeof(double) + ((tot + 1) / 2) * sizeof(uint32_t);
    b = (b + 15) & ~(size_t)15;
    b += (size_t)16 * kNoff * sizeof(uint32_t);
    return b;
}

static uint64_t max_points(const vgx_config* c, const IterGeom& g) {
    if (!g.adaptive_sampler) return (uint64_t)c->nhcube * (uint64_t)g.avg_nh;
    const double adapt = c->neval_frac * (double)c->neval;
    return (uint64_t)(adapt * (1.0 + 1e-9)) + 2 + (uint64_t)c->nhcube * (uint64_t)g.nmin;
}

struct Run {
    IterGeom g;
    FusedPtrs io;
    StradArgs sa;
    int fused_blocks;
    int strad_blocks;
    size_t lds;
    int nblk_alloc;
    uint64_t cap;
    uint32_t maxtiles;
    Scalars* sc;
    uint64_t* stat;
    int32_t* nh;
    float* fparams;
    AxisC* axes;
    double* errv;
    double* sigf_part;
    int hist_tot;  // dim * nb when the fused sampler trains (records and rows)
    int exact_train;  // training histogram summed in point order from the records
    double* rows_s;   // fast mode: [kTrainBlocks][dim * nb] rows of the training histogram
    uint32_t* rows_c;
};

static int prepare(vgx_ctx* ctx, const vgx_config* c, vgx_state* st, Run* R) {
    int rc = build_geom(c, ctx->info, &R->g);
    if (rc) return rc;
    const IterGeom& g = R->g;
    R->cap = max_points(c, g);
    if (R->cap >= 0xFFFFFFF0ull) return VGX_ERANGE;
    R->maxtiles = (uint32_t)((R->cap + kT - 1) / kT) + 1;
    R->nblk_alloc = (int)((c->nhcube + kAllocPerBlock - 1) / kAllocPerBlock);
    R->fused_blocks = ctx->info.cus;
    R->strad_blocks = (int)((2ull * R->maxtiles + kStradThreads - 1) / kStradThreads);
    R->lds = fused_lds(g);
    if (R->lds > ctx->info.max_lds) return VGX_ERANGE;
    if ((rc = ctx->nh.ensure((size_t)c->nhcube * 4))) return rc;
    if ((rc = ctx->off.ensure(((size_t)c->nhcube + 1) * 4))) return rc;
    if ((rc = ctx->blocksum.ensure((size_t)R->nblk_alloc * 3 * 8))) return rc;
    if ((rc = ctx->tile_h.ensure((size_t)R->maxtiles * 4))) return rc;
    if ((rc = ctx->tile_slots.ensure((size_t)R->maxtiles * 2 * sizeof(Slot)))) return rc;
    if ((rc = ctx->tile_part.ensure((size_t)R->maxtiles * 3 * sizeof(double)))) return rc;
    if ((rc = ctx->blk_part.ensure((size_t)R->fused_blocks * 6 * 8))) return rc;
    if ((rc = ctx->strad_part.ensure((size_t)R->strad_blocks * 6 * 8))) return rc;
    R->errv = nullptr;
    if (g.train_errors) {
        if ((rc = ctx->errv.ensure((size_t)c->nhcube * 8))) return rc;
        R->errv = (double*)ctx->errv.ptr;
    }
    // misc: Scalars | axes | sigf partials | float integrand parameters
    const size_t misc = 1024 + VGX_MAXDIM * sizeof(AxisC) + 2 * 256 * sizeof(double) + 1024 * 1024;
    if ((rc = ctx->misc.ensure(misc))) return rc;
    R->sc = (Scalars*)ctx->misc.ptr;
    R->axes = (AxisC*)((char*)ctx->misc.ptr + 1024);
    R->sigf_part = (double*)((char*)R->axes + VGX_MAXDIM * sizeof(AxisC));
    R->fparams = (float*)((char*)R->sigf_part + 2 * 256 * sizeof(double));
    R->stat = (uint64_t*)ctx->blocksum.ptr;
    R->nh = (int32_t*)ctx->nh.ptr;
    FusedPtrs& io = R->io;
    io.axes = R->axes;
    io.grid = st->grid;
    io.off = (const uint32_t*)ctx->off.ptr;
    io.tile_h = (const int32_t*)ctx->tile_h.ptr;
    io.sigf = st->sigf;
    io.slots = (Slot*)ctx->tile_slots.ptr;
    io.tile_part = (double*)ctx->tile_part.ptr;
    io.blk_part = (double*)ctx->blk_part.ptr;
    io.sum_f = st->sum_f;
    io.n_f = st->n_f;
    io.rec_bins = nullptr;
    io.rec_c2 = nullptr;
    io.rec_stride = 0;
    io.hpart_s = nullptr;
    io.hpart_c = nullptr;
    R->hist_tot = 0;
    R->exact_train = (c->flags & VGX_FLAG_EXACT_ADAPT) ? 1 : 0;
    R->rows_s = nullptr;
    R->rows_c = nullptr;
    if (g.train_points) {
        // one record per point of a batch (at most cap points): bins, then
        // (w f n)^2; rows of the fast-mode histogram per block
        R->hist_tot = c->dim * g.nb;
        // rows padded to whole 256-point tiles (the ordered histogram reads a
        // lane's 4 records with one 16-byte load, also in a partial tile)
        const size_t stride = (R->cap + kT - 1) / kT * kT;
        const size_t wbytes = stride * (size_t)((c->dim + 1) / 2) * sizeof(uint32_t);
        if ((rc = ctx->trec.ensure(wbytes + stride * sizeof(double)))) return rc;
        io.rec_bins = (uint32_t*)ctx->trec.ptr;
        io.rec_c2 = (double*)((char*)ctx->trec.ptr + wbytes);
        io.rec_stride = (uint32_t)stride;
        if (g.hist_lds) {
            // one row per fused workgroup
            const size_t cells = (size_t)R->fused_blocks * (size_t)R->hist_tot;
            if ((rc = ctx->hist_part.ensure(cells * (sizeof(double) + sizeof(uint32_t))))) return rc;
            io.hpart_s = (double*)ctx->hist_part.ptr;
            io.hpart_c = (uint32_t*)((double*)ctx->hist_part.ptr + cells);
        } else if (!R->exact_train) {
            const size_t cells = (size_t)kTrainBlocks * (size_t)R->hist_tot;
            if ((rc = ctx->hist_part.ensure(cells * (sizeof(double) + sizeof(uint32_t))))) return rc;
            R->rows_s = (double*)ctx->hist_part.ptr;
            R->rows_c = (uint32_t*)((double*)ctx->hist_part.ptr + cells);
        }
    }
    io.errv = R->errv;
    io.sc = R->sc;
    StradArgs& sa = R->sa;
    sa.errv = R->errv;
    sa.off = io.off;
    sa.slots = io.slots;
    sa.sigf = st->sigf;
    sa.strad_part = (double*)ctx->strad_part.ptr;
    sa.cont = nullptr;
    sa.sc = R->sc;
    return VGX_OK;
}

static void set_iter(Run* R, const vgx_config* c, int it) {
    R->g.iteration = c->iteration + (uint32_t)it;
    R->g.itn = (uint32_t)it;
}

static int init_axes(Run* R, hipStream_t s) {
    hipLaunchKernelGGL(axes_init_kernel, dim3(1), dim3(64), 0, s, R->g, R->axes);
    return check_launch();
}

static int run_init(Run* R, vgx_state* st, hipStream_t s, int new_run) {
    const IterGeom& g = R->g;
    const int tot = g.dim * g.nb;
    const int nb = (tot + 255) / 256 < 64 ? (tot + 255) / 256 : 64;
    hipLaunchKernelGGL(iter_init_kernel, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, g, st->sum_f, st->n_f,
                       R->sc, new_run);
    return check_launch();
}

static int alloc_kernels(vgx_ctx* ctx, Run* R, vgx_state* st, hipStream_t s) {
    const IterGeom& g = R->g;
    hipLaunchKernelGGL(alloc_kernel, dim3(R->nblk_alloc), dim3(kAllocThreads), 0, s, g, st->sigf,
                       (const double*)st->sum_sigf, R->nh, R->stat, (const Scalars*)R->sc, st->sum_f, st->n_f);
    hipLaunchKernelGGL(offsets_kernel, dim3(R->nblk_alloc), dim3(kAllocThreads), 0, s, g,
                       (const int32_t*)R->nh, (const uint64_t*)R->stat, R->nblk_alloc,
                       (uint32_t*)ctx->off.ptr, (int32_t*)ctx->tile_h.ptr, R->sc, R->cap);
    return check_launch();
}

// Fused sampler + straddlers for the current batch.  reduce = 1 (staged
// batches) adds the batch's partials into the iteration sums at once;
// otherwise finalize_kernel reads them.
template <class Fcn>
static int accumulate_kernels(vgx_ctx* ctx, Run* R, const Fcn& fcn, hipStream_t s, int timed, int reduce) {
    const IterGeom& g = R->g;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool rec = timed && timed_slot(ctx, kTimedFused, &e0, &e1);
    int rc = rec ? launch_fused(g, R->io, fcn, R->fused_blocks, R->lds, s, e0, e1)
                 : launch_fused(g, R->io, fcn, R->fused_blocks, R->lds, s);
    if (rc) return rc;
    if (R->io.hpart_s) {
        // the fused launch's workgroup rows, summed in workgroup order
        const HistArgs h{R->io.hpart_s, R->io.hpart_c, R->fused_blocks, R->hist_tot, R->io.sum_f, R->io.n_f};
        const int nh = (R->hist_tot + kHrCols - 1) / kHrCols;
        hipLaunchKernelGGL(strad_hist_kernel, dim3(R->strad_blocks + nh), dim3(kStradThreads), 0, s, g, R->sa, h,
                           R->strad_blocks);
    } else if (R->io.rec_c2 && R->exact_train) {
        // the reference's order: every bin summed over the batch's points in order
        rc = train_hist_records(ctx, g.dim, g.ninc, g.nb, R->io.rec_bins, (int64_t)R->io.rec_stride, R->io.rec_c2,
                                &R->sc->b_p0, (int64_t)R->cap, R->io.sum_f, R->io.n_f, s);
        if (rc) return rc;
        hipLaunchKernelGGL(strad_kernel, dim3(R->strad_blocks), dim3(kStradThreads), 0, s, g, R->sa);
    } else if (R->io.rec_c2) {
        const size_t lds = (size_t)g.nb * (sizeof(double) + sizeof(uint32_t));
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (timed && timed_slot(ctx, kTimedTrainRows, &t0, &t1))
            hipExtLaunchKernelGGL(train_rows_kernel, dim3(kTrainBlocks * g.dim), dim3(64), (uint32_t)lds, s, t0, t1,
                                  0, g, (const uint32_t*)R->io.rec_bins, (const double*)R->io.rec_c2,
                                  R->io.rec_stride, R->rows_s, R->rows_c, (const Scalars*)R->sc);
        else
            hipLaunchKernelGGL(train_rows_kernel, dim3(kTrainBlocks * g.dim), dim3(64), lds, s, g,
                               (const uint32_t*)R->io.rec_bins, (const double*)R->io.rec_c2, R->io.rec_stride,
                               R->rows_s, R->rows_c, (const Scalars*)R->sc);
        const HistArgs h{R->rows_s, R->rows_c, kTrainBlocks, R->hist_tot, R->io.sum_f, R->io.n_f};
        const int nh = (R->hist_tot + kHrCols - 1) / kHrCols;
        hipLaunchKernelGGL(strad_hist_kernel, dim3(R->strad_blocks + nh), dim3(kStradThreads), 0, s, g, R->sa, h,
                           R->strad_blocks);
    } else {
        hipLaunchKernelGGL(strad_kernel, dim3(R->strad_blocks), dim3(kStradThreads), 0, s, g, R->sa);
    }
    if (reduce)
        hipLaunchKernelGGL(batch_reduce_kernel, dim3(1), dim3(kReduceThreads), 0, s,
                           (const double*)R->io.blk_part, R->fused_blocks, (const double*)R->sa.strad_part,
                           R->strad_blocks, R->sc);
    return check_launch();
}

// Iteration close.  parts = 1 when the iteration was one batch whose
// partials are still unreduced (vgx_integrate).
static int iter_end_kernels(vgx_ctx* ctx, Run* R, const vgx_config* c, vgx_state* st, hipStream_t s,
                            int parts) {
    const IterGeom& g = R->g;
    const int has_sigf = c->beta >= 0 && c->nhcube > 1 && !c->adapt_to_errors;
    if (g.adaptive_strat && !g.fast_pow)
        hipLaunchKernelGGL(sigf_pow_kernel, dim3(256), dim3(256), 0, s, g, st->sigf, R->sigf_part,
                           (const Scalars*)R->sc);
    if (g.train_errors) {
        const unsigned nb = (unsigned)((c->nhcube + 255) / 256 < 4096 ? (c->nhcube + 255) / 256 : 4096);
        if (g.rng32)
            hipLaunchKernelGGL(errors_train_kernel<true>, dim3(nb), dim3(256), 0, s, g, (const AxisC*)R->axes,
                               (const double*)st->grid, (const uint32_t*)ctx->off.ptr, (const double*)R->errv,
                               st->sum_f, st->n_f, (const Scalars*)R->sc);
        else
            hipLaunchKernelGGL(errors_train_kernel<false>, dim3(nb), dim3(256), 0, s, g, (const AxisC*)R->axes,
                               (const double*)st->grid, (const uint32_t*)ctx->off.ptr, (const double*)R->errv,
                               st->sum_f, st->n_f, (const Scalars*)R->sc);
    }
    const int do_adapt = c->alpha > 0 && c->adapt;
    const int has_training = g.train_points || g.train_errors;
    const int exact = (c->flags & VGX_FLAG_EXACT_ADAPT) ? 1 : 0;
    int rc;
    if (do_adapt) {
        AdaptArgs a;
        int parallel = 0;
        rc = adapt_in_place_setup(ctx, c->dim, c->ninc, st->grid, c->ld, st->sum_f, st->n_f, has_training,
                                  c->alpha, exact, &R->sc->status, &a, &parallel);
        if (rc) return rc;
        if (parallel) {
            hipLaunchKernelGGL(close_kernel, dim3(c->dim + 1), dim3(1024), 0, s, g, R->sc, *st, has_sigf,
                               (const double*)R->sigf_part, 256, (const double*)R->io.blk_part,
                               parts ? R->fused_blocks : 0, (const double*)R->sa.strad_part,
                               parts ? R->strad_blocks : 0, a);
            return check_launch();
        }
    }
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kReduceThreads), 0, s, g, R->sc, *st, has_sigf,
                       (const double*)R->sigf_part, 256, (const double*)R->io.blk_part,
                       parts ? R->fused_blocks : 0, (const double*)R->sa.strad_part,
                       parts ? R->strad_blocks : 0);
    rc = check_launch();
    if (rc) return rc;
    if (do_adapt)
        rc = adapt_in_place(ctx, c->dim, c->ninc, st->grid, c->ld, st->sum_f, st->n_f, has_training,
                            c->alpha, exact, &R->sc->status, s);
    return rc;
}

static int run_end(Run* R, const vgx_config* c, vgx_state* st, hipStream_t s) {
    hipLaunchKernelGGL(sigf_reset_kernel, dim3(256), dim3(256), 0, s, st->sigf, (int64_t)c->nhcube, R->sc);
    return check_launch();
}

static int make_fcn_and_run(vgx_ctx* ctx, Run* R, const vgx_config* c, const vgx_integrand* f,
                            hipStream_t s, int timed, int convert, int reduce) {
    const bool f32 = c->precision == VGX_PREC_F32;
    if (f->kind == VGX_FCN_BUFFER) {
        if (f32) {
            BufferFcn<true> b{f->fbuf, f->fbase};
            return accumulate_kernels(ctx, R, b, s, timed, reduce);
        }
        BufferFcn<false> b{f->fbuf, f->fbase};
        return accumulate_kernels(ctx, R, b, s, timed, reduce);
    }
    if (f->kind == VGX_FCN_GAUSSMIX || f->kind == VGX_FCN_DIAGRIDGE) {
        const int nc = f->ncomp;
        if (nc < 1 || !f->A || !f->B || !f->c) return VGX_EINVAL;
        const int ncc = f->kind == VGX_FCN_GAUSSMIX ? nc * c->dim : nc;
        if (f32) {
            if ((size_t)(2 * nc + ncc) * 4 > 1024 * 1024) return VGX_ERANGE;
            float* A = R->fparams;
            float* B = A + nc;
            float* C = B + nc;
            if (convert) {
                hipLaunchKernelGGL(to_float_kernel, dim3(64), dim3(256), 0, s, f->A, A, nc);
                hipLaunchKernelGGL(to_float_kernel, dim3(64), dim3(256), 0, s, f->B, B, nc);
                hipLaunchKernelGGL(to_float_kernel, dim3(64), dim3(256), 0, s, f->c, C, ncc);
            }
            if (f->kind == VGX_FCN_GAUSSMIX) {
                GaussMix32 fn{nc, A, B, C};
                return accumulate_kernels(ctx, R, fn, s, timed, reduce);
            }
            DiagRidge32 fn{nc, A, B, C};
            return accumulate_kernels(ctx, R, fn, s, timed, reduce);
        }
        if (f->kind == VGX_FCN_GAUSSMIX) {
            GaussMix64 fn{nc, f->A, f->B, f->c};
            return accumulate_kernels(ctx, R, fn, s, timed, reduce);
        }
        DiagRidge64 fn{nc, f->A, f->B, f->c};
        return accumulate_kernels(ctx, R, fn, s, timed, reduce);
    }
    return VGX_EINVAL;
}

}  // namespace vgx

using namespace vgx;

extern "C" int vgx_integrate(vgx_ctx* ctx, const vgx_config* c, const vgx_integrand* f, vgx_state* st,
                             void* stream) {
    if (!ctx || !c || !f || !st) return VGX_EINVAL;
    if (f->kind == VGX_FCN_BUFFER || c->uniforms) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    Run R;
    int rc = prepare(ctx, c, st, &R);
    if (rc) return rc;
    if ((rc = init_axes(&R, s))) return rc;
    if ((rc = run_init(&R, st, s, 1))) return rc;
    for (int it = 0; it < c->nitn; ++it) {
        set_iter(&R, c, it);
        if ((rc = alloc_kernels(ctx, &R, st, s))) return rc;
        if ((rc = make_fcn_and_run(ctx, &R, c, f, s, 1, it == 0, 0))) return rc;
        if ((rc = iter_end_kernels(ctx, &R, c, st, s, 1))) return rc;
    }
    return run_end(&R, c, st, s);
}

extern "C" int vgx_iter_begin(vgx_ctx* ctx, const vgx_config* c, vgx_state* st, int it, int32_t* n_h,
                              int64_t* total_host, void* stream) {
    if (!ctx || !c || !st) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    Run R;
    int rc = prepare(ctx, c, st, &R);
    if (rc) return rc;
    if ((rc = init_axes(&R, s))) return rc;
    set_iter(&R, c, it);
    if ((rc = run_init(&R, st, s, 1))) return rc;
    if ((rc = alloc_kernels(ctx, &R, st, s))) return rc;
    if (n_h) VGX_HIP(hipMemcpyAsync(n_h, R.nh, (size_t)c->nhcube * 4, hipMemcpyDeviceToDevice, s));
    if (total_host) {
        Scalars tmp;
        VGX_HIP(hipMemcpyAsync(&tmp, R.sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
        VGX_HIP(hipStreamSynchronize(s));
        if (tmp.overflow) return VGX_ERANGE;
        *total_host = (int64_t)tmp.total;
    }
    return VGX_OK;
}

extern "C" int vgx_iter_sample(vgx_ctx* ctx, const vgx_config* c, vgx_state* st, int it, int64_t h0,
                               int64_t h1, int64_t p0, int64_t p1, double* x, double* y, double* wgt,
                               int64_t* hcube, void* stream) {
    if (!ctx || !c || !st || !x) return VGX_EINVAL;
    if (h0 < 0 || h1 > c->nhcube || h0 >= h1 || p0 < 0 || p1 <= p0) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    Run R;
    int rc = prepare(ctx, c, st, &R);
    if (rc) return rc;
    if ((rc = init_axes(&R, s))) return rc;
    set_iter(&R, c, it);
    const int64_t n = p1 - p0;
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (R.g.rng32)
        hipLaunchKernelGGL(sample_kernel<true>, dim3((unsigned)b), dim3(256), 0, s, R.g, (const AxisC*)R.axes,
                           (const double*)st->grid, (const uint32_t*)ctx->off.ptr, (uint32_t)h0, (uint32_t)h1,
                           (uint32_t)p0, (uint32_t)p1, x, y, wgt, hcube, (const Scalars*)R.sc);
    else
        hipLaunchKernelGGL(sample_kernel<false>, dim3((unsigned)b), dim3(256), 0, s, R.g, (const AxisC*)R.axes,
                           (const double*)st->grid, (const uint32_t*)ctx->off.ptr, (uint32_t)h0, (uint32_t)h1,
                           (uint32_t)p0, (uint32_t)p1, x, y, wgt, hcube, (const Scalars*)R.sc);
    return check_launch();
}

extern "C" int vgx_iter_accumulate(vgx_ctx* ctx, const vgx_config* c, vgx_state* st, int it, int64_t h0,
                                   int64_t h1, int64_t p0, int64_t p1, const double* fv, void* stream) {
    if (!ctx || !c || !st || !fv) return VGX_EINVAL;
    if (h0 < 0 || h1 > c->nhcube || h0 >= h1 || p0 < 0 || p1 <= p0) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    Run R;
    int rc = prepare(ctx, c, st, &R);
    if (rc) return rc;
    if ((rc = init_axes(&R, s))) return rc;
    set_iter(&R, c, it);
    const uint32_t nhb = (uint32_t)(h1 - h0);
    hipLaunchKernelGGL(batch_tiles_kernel, dim3((nhb + 255) / 256 < 1024 ? (nhb + 255) / 256 : 1024),
                       dim3(256), 0, s, (const uint32_t*)ctx->off.ptr, (int32_t*)ctx->tile_h.ptr, R.sc,
                       (uint32_t)h0, (uint32_t)h1, (uint32_t)p0, (uint32_t)p1);
    if ((rc = check_launch())) return rc;
    vgx_integrand f{};
    f.kind = VGX_FCN_BUFFER;
    f.fbuf = fv;
    f.fbase = p0;
    return make_fcn_and_run(ctx, &R, c, &f, s, 0, 0, 1);
}

extern "C" int vgx_iter_end(vgx_ctx* ctx, const vgx_config* c, vgx_state* st, int it, void* stream) {
    if (!ctx || !c || !st) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    Run R;
    int rc = prepare(ctx, c, st, &R);
    if (rc) return rc;
    if ((rc = init_axes(&R, s))) return rc;
    set_iter(&R, c, it);
    if ((rc = iter_end_kernels(ctx, &R, c, st, s, 0))) return rc;
    return run_end(&R, c, st, s);
}
