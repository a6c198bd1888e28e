/*
 * vegas_oracle.c -- FP64 CPU oracle for the vegas+ hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see vegas_oracle.h): a plain-C restatement of
 * the behaviour of gplepage/vegas 6.4.1, src/vegas/_vegas.pyx (pyx:N).  It
 * exists to check the HIP kernels and to time a CPU baseline; the product
 * never calls it.  Build: -O2 -ffp-contract=off (oracle/Makefile).
 */
#include "vegas_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* pyx:34-36: TINY = 10**(min_10_exp+50), HUGE = 10**(max_10_exp-50),
 * EPSILON = DBL_EPSILON * 1e4 */
static const double TINY = 1e-257;
static const double HUGE_ = 1e258;
#define EPSILON_ (DBL_EPSILON * 1e4)

double vgo_tiny(void) { return TINY; }
double vgo_huge(void) { return HUGE_; }
double vgo_epsilon(void) { return EPSILON_; }

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random numbers:
 * as easy as 1, 2, 3").  Round: (hi0,lo0)=M0*c0, (hi1,lo1)=M1*c2,
 * c' = (hi1^c1^k0, lo1, hi0^c3^k1, lo0); key bumped by the Weyl constants. */
void vgo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void block_words(uint64_t seed, uint32_t it, int64_t p, uint32_t blk, uint32_t w[4]) {
    uint32_t ctr[4] = {(uint32_t)((uint64_t)p & 0xffffffffu), (uint32_t)((uint64_t)p >> 32), it, blk};
    uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    vgo_philox4x32_10(ctr, key, w);
}

void vgo_uniforms_f64(uint64_t seed, uint32_t it, int64_t p0, int64_t n, int dim, double* U) {
    for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; 2 * j < dim; ++j) {
            uint32_t w[4];
            block_words(seed, it, p0 + i, (uint32_t)j, w);
            for (int e = 0; e < 2 && 2 * j + e < dim; ++e) {
                double a = (double)(w[2 * e] >> 5), b = (double)(w[2 * e + 1] >> 6);
                U[i * dim + 2 * j + e] = (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
            }
        }
    }
}

void vgo_uniforms_f32(uint64_t seed, uint32_t it, int64_t p0, int64_t n, int dim, double* U) {
    for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; 4 * j < dim; ++j) {
            uint32_t w[4];
            block_words(seed, it, p0 + i, (uint32_t)j, w);
            for (int e = 0; e < 4 && 4 * j + e < dim; ++e)
                U[i * dim + 4 * j + e] = (double)(w[e] >> 8) * (1.0 / 16777216.0);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* AdaptiveMap.make_uniform (pyx:205-241) via numpy.linspace:
 * node i = i*step + lo with step = (hi-lo)/n, last node = hi exactly.
 * Padding beyond ninc repeats the last node (inc padding 0). */
void vgo_make_uniform(int dim, const double* lo, const double* hi, const int64_t* ninc,
                      int ld, double* grid, double* inc) {
    for (int d = 0; d < dim; ++d) {
        int64_t n = ninc[d];
        double step = (hi[d] - lo[d]) / (double)n;
        for (int64_t i = 0; i < n; ++i) grid[d * ld + i] = (double)i * step + lo[d];
        grid[d * ld + n] = hi[d];
        for (int64_t i = n + 1; i < ld; ++i) grid[d * ld + i] = hi[d];
        for (int64_t i = 0; i < n; ++i)
            inc[d * (ld - 1) + i] = grid[d * ld + i + 1] - grid[d * ld + i];
        for (int64_t i = n; i < ld - 1; ++i) inc[d * (ld - 1) + i] = 0.0;
    }
}

/* AdaptiveMap.map (pyx:310-360) */
int vgo_map(int dim, int ld, const double* grid, const double* inc, const int64_t* ninc,
            const double* y, double* x, double* jac, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        jac[i] = 1.;
        for (int d = 0; d < dim; ++d) {
            int64_t ni = ninc[d];
            double y_ninc = y[i * dim + d] * (double)ni;
            int iy = (int)floor(y_ninc);
            double dy_ninc = y_ninc - iy;
            if (iy < ni) {
                x[i * dim + d] = grid[d * ld + iy] + inc[d * (ld - 1) + iy] * dy_ninc;
                jac[i] *= inc[d * (ld - 1) + iy] * (double)ni;
            } else {
                x[i * dim + d] = grid[d * ld + ni];
                jac[i] *= inc[d * (ld - 1) + ni - 1] * (double)ni;
            }
        }
    }
    return 0;
}

/* AdaptiveMap.jac1d (pyx:265-295) */
void vgo_jac1d(int dim, int ld, const double* inc, const int64_t* ninc, const double* y,
               double* jac, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < dim; ++d) {
            int64_t ni = ninc[d];
            double y_ninc = y[i * dim + d] * (double)ni;
            int iy = (int)floor(y_ninc);
            jac[i * dim + d] = (iy < ni) ? inc[d * (ld - 1) + iy] * (double)ni
                                         : inc[d * (ld - 1) + ni - 1] * (double)ni;
        }
}

/* numpy.searchsorted(side='right') over the live nodes grid[0..ni] */
static int64_t search_right(const double* g, int64_t len, double v) {
    int64_t lo = 0, hi = len;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (g[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* AdaptiveMap.invmap (pyx:362-416) */
int vgo_invmap(int dim, int ld, const double* grid, const double* inc, const int64_t* ninc,
               const double* x, double* y, double* jac, int64_t n) {
    for (int64_t i = 0; i < n; ++i) jac[i] = 1.;
    for (int d = 0; d < dim; ++d) {
        int64_t ni = ninc[d];
        const double* g = grid + d * ld;
        const double* in = inc + d * (ld - 1);
        for (int64_t i = 0; i < n; ++i) {
            int64_t iy = search_right(g, ni + 1, x[i * dim + d]);
            if (iy > 0 && iy <= ni) {
                int64_t k = iy - 1;
                y[i * dim + d] = ((double)k + (x[i * dim + d] - g[k]) / in[k]) / (double)ni;
                jac[i] *= in[k] * (double)ni;
            } else if (iy <= 0) {
                y[i * dim + d] = 0.;
                jac[i] *= in[0] * (double)ni;
            } else {
                y[i * dim + d] = 1.0;
                jac[i] *= in[ni - 1] * (double)ni;
            }
        }
    }
    return 0;
}

/* AdaptiveMap.add_training_data (pyx:421-464): axis-major, point order */
void vgo_add_training_data(int dim, int ld, const int64_t* ninc, const double* y,
                           const double* f, int64_t n, double* sum_f, double* n_f) {
    for (int d = 0; d < dim; ++d) {
        int64_t ni = ninc[d];
        for (int64_t i = 0; i < n; ++i) {
            double yi = y[i * dim + d];
            if (yi > 0 && yi < 1) {
                int iy = (int)floor(yi * (double)ni);
                sum_f[d * (ld - 1) + iy] += fabs(f[i]);
                n_f[d * (ld - 1) + iy] += 1;
            }
        }
    }
}

/* AdaptiveMap.adapt (pyx:467-594).  avg_f is one buffer shared by all axes,
 * initialised to ones (pyx:534); values written for one axis persist into
 * the next when that axis has no training data. */
int vgo_adapt(int dim, int ld, const double* grid, const double* inc, const int64_t* ninc,
              int has_training, const double* sum_f, const double* n_f, double alpha,
              const int64_t* new_ninc, int new_ld, double* new_grid, double* new_inc) {
    int64_t max_new = 0, min_new = INT64_MAX, max_old = 0;
    for (int d = 0; d < dim; ++d) {
        if (new_ninc[d] > max_new) max_new = new_ninc[d];
        if (new_ninc[d] < min_new) min_new = new_ninc[d];
        if (ninc[d] > max_old) max_old = ninc[d];
    }
    if (min_new < 1) return -1;
    for (int64_t i = 0; i < (int64_t)dim * new_ld; ++i) new_grid[i] = NAN;
    if (max_new == 1) {
        for (int d = 0; d < dim; ++d) {
            new_grid[d * new_ld + 0] = grid[d * ld + 0];
            new_grid[d * new_ld + 1] = grid[d * ld + ninc[d]];
            new_inc[d * (new_ld - 1)] = new_grid[d * new_ld + 1] - new_grid[d * new_ld + 0];
        }
        return 0;
    }
    int64_t nbuf = ld - 1;
    double* avg_f = (double*)malloc(sizeof(double) * (size_t)nbuf);
    double* tmp_f = (double*)malloc(sizeof(double) * (size_t)nbuf);
    for (int64_t i = 0; i < nbuf; ++i) avg_f[i] = 1.0;
    for (int d = 0; d < dim; ++d) {
        int64_t old = ninc[d];
        int64_t nn = new_ninc[d];
        const double* g = grid + d * ld;
        const double* in = inc + d * (ld - 1);
        double* ng = new_grid + d * new_ld;
        if (alpha != 0 && old > 1) {
            if (has_training) {
                for (int64_t i = 0; i < old; ++i) {
                    double nf = n_f[d * (ld - 1) + i];
                    avg_f[i] = (nf > 0) ? sum_f[d * (ld - 1) + i] / nf : 0.;
                }
            }
            if (alpha > 0) {
                tmp_f[0] = fabs(7. * avg_f[0] + avg_f[1]) / 8.;
                tmp_f[old - 1] = fabs(7. * avg_f[old - 1] + avg_f[old - 2]) / 8.;
                double s = tmp_f[0] + tmp_f[old - 1];
                for (int64_t i = 1; i < old - 1; ++i) {
                    tmp_f[i] = fabs(6. * avg_f[i] + avg_f[i - 1] + avg_f[i + 1]) / 8.;
                    s += tmp_f[i];
                }
                if (s > 0) {
                    for (int64_t i = 0; i < old; ++i) avg_f[i] = tmp_f[i] / s + TINY;
                } else {
                    for (int64_t i = 0; i < old; ++i) avg_f[i] = TINY;
                }
                for (int64_t i = 0; i < old; ++i)
                    if (avg_f[i] > 0 && avg_f[i] <= 0.99999999)
                        avg_f[i] = pow(-(1 - avg_f[i]) / log(avg_f[i]), alpha);
            }
        }
        ng[0] = g[0];
        ng[nn] = g[old];
        double acc = 0., f_ninc = 0.;
        int64_t j = -1;
        for (int64_t i = 0; i < old; ++i) f_ninc += avg_f[i];
        f_ninc /= (double)nn;
        for (int64_t i = 1; i < nn; ++i) {
            int stopped = 0;
            while (acc < f_ninc) {
                j += 1;
                if (j < old) acc += avg_f[j];
                else { stopped = 1; break; }
            }
            if (stopped) break;
            acc -= f_ninc;
            ng[i] = g[j + 1] - (acc / avg_f[j]) * in[j];
        }
        for (int64_t i = nn + 1; i < new_ld; ++i) ng[i] = ng[nn];
        for (int64_t i = 0; i < nn; ++i) new_inc[d * (new_ld - 1) + i] = ng[i + 1] - ng[i];
        for (int64_t i = nn; i < new_ld - 1; ++i) new_inc[d * (new_ld - 1) + i] = 0.0;
    }
    free(avg_f);
    free(tmp_f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Integrator.set geometry (pyx:1321-1446), integer logic with float ** and log */
void vgo_geometry(int dim, int64_t neval, int neval_given, const int64_t* nstrat_in,
                  double neval_frac, double beta, int adapt_to_errors, int uniform_nstrat,
                  int64_t maxinc_axis, double max_mem, int minimize_mem,
                  int64_t min_neval_batch, int64_t max_neval_hcube, vgo_geometry_out* out) {
    memset(out, 0, sizeof(*out));
    double nf = (beta == 0 || adapt_to_errors) ? 0.0 : neval_frac;
    int64_t* ns_ = out->nstrat;
    if (neval_frac < 0 || neval_frac >= 1) { out->status = -1; return; }
    if (nstrat_in) {
        int64_t nh = 1;
        for (int d = 0; d < dim; ++d) {
            if (nstrat_in[d] < 1) { out->status = -1; return; }
            ns_[d] = nstrat_in[d];
            nh *= ns_[d];
        }
        if (!neval_given) {
            neval = (int64_t)(2. * (double)nh / (1. - nf));
        } else if ((double)neval < 2. * (double)nh / (1. - nf)) {
            out->status = -1;
            return;
        }
    } else {
        if (neval < 2) { out->status = -1; return; }
        int64_t ns = (int64_t)pow(fabs((1 - nf) * (double)neval / 2.), 1. / dim);
        if (ns < 1) ns = 1;
        int64_t dd = (int64_t)((log((1 - nf) * (double)neval / 2.) - dim * log((double)ns)) /
                               log(1 + 1. / (double)ns));
        double nh = pow((double)(ns + 1), (double)dd) * pow((double)ns, (double)(dim - dd));
        if (nh > max_mem && !minimize_mem) { out->status = -2; return; }
        if (uniform_nstrat) dd = 0;
        if (dd < 0) dd = dim + dd > 0 ? dim + dd : 0; /* python slice nstrat[:d] */
        if (dd > dim) dd = dim;
        for (int d = 0; d < dim; ++d) ns_[d] = (d < dd) ? ns + 1 : ns;
    }
    if (adapt_to_errors) {
        for (int d = 0; d < dim; ++d) out->ninc[d] = ns_[d];
    } else {
        int64_t ni = (int64_t)((double)neval / 10.);
        if (ni > maxinc_axis) ni = maxinc_axis;
        for (int d = 0; d < dim; ++d) {
            if (ni >= ns_[d]) {
                out->ninc[d] = (int64_t)((double)ni / (double)ns_[d]) * ns_[d];
            } else if (ns_[d] <= maxinc_axis) {
                out->ninc[d] = ns_[d];
            } else {
                ns_[d] = (int64_t)((double)ns_[d] / (double)ni) * ni;
                out->ninc[d] = ni;
            }
        }
    }
    int64_t nh = 1;
    for (int d = 0; d < dim; ++d) nh *= ns_[d];
    out->nhcube = nh;
    out->neval = neval;
    if (nh == 1) out->min_neval_hcube = neval;
    else out->min_neval_hcube = (int64_t)((1 - nf) * (double)neval / (double)nh);
    if (out->min_neval_hcube < 2) out->min_neval_hcube = 2;
    int64_t ws = min_neval_batch + max_neval_hcube;
    if (ws > neval) ws = neval + 1;
    out->workspace = ws;
    if ((double)((3 * dim + 3) * ws + (minimize_mem ? 0 : nh)) > max_mem) { out->status = -2; return; }
    out->status = 0;
}

/* ------------------------------------------------------------------------ */
int vgo_gaussmix_eval(void* ctx, const double* x, int64_t n, int dim, double* f) {
    const vgo_gaussmix* g = (const vgo_gaussmix*)ctx;
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int k = 0; k < g->ncomp; ++k) {
            double r2 = 0.0;
            for (int d = 0; d < dim; ++d) {
                double t = x[i * dim + d] - g->c[k * dim + d];
                r2 += t * t;
            }
            acc += g->A[k] * exp(-g->B[k] * r2);
        }
        f[i] = acc;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* One iteration: sampler (pyx:1650-1765) + accumulation (pyx:2096-2197). */
int vgo_iterate(const vgo_iter_params* P, const double* grid, const double* inc,
                const int64_t* ninc, double* sigf, double sum_sigf_in, double* sum_f,
                double* n_f, vgo_fcn_t fcn, void* ctx, vgo_iter_out* out) {
    const int dim = P->dim;
    const int64_t nhcube = P->nhcube;
    const double dv_y = 1. / (double)nhcube;
    const double neval_sigf = (P->beta > 0 && sum_sigf_in > 0 && !P->adapt_to_errors)
                                  ? P->neval_frac * (double)P->neval / sum_sigf_in
                                  : 0.0;
    const int64_t avg_nh = (int64_t)((double)P->neval / (double)nhcube);
    const int64_t nmin = P->min_neval_hcube;
    const int64_t nmax = P->max_neval_hcube > nmin ? P->max_neval_hcube : nmin;
    const int adaptive_sampler = P->beta > 0 && nhcube > 1 && !P->adapt_to_errors;
    const int adaptive_strat = P->beta > 0 && nhcube > 1 && P->adapt && !P->adapt_to_errors;
    const int train_points = !P->adapt_to_errors && P->adapt && P->alpha > 0;
    const double eps = EPSILON_;
    const int64_t max_nh_batch = P->min_neval_batch / 2 + 1;

    int64_t cap = P->min_neval_batch + nmax + 16;
    double* y = (double*)malloc(sizeof(double) * (size_t)(cap * dim));
    double* x = (double*)malloc(sizeof(double) * (size_t)(cap * dim));
    double* U = (double*)malloc(sizeof(double) * (size_t)(cap * dim));
    double* jac = (double*)malloc(sizeof(double) * (size_t)cap);
    double* fdv2 = (double*)malloc(sizeof(double) * (size_t)cap);
    double* fx = (double*)malloc(sizeof(double) * (size_t)cap);
    int64_t* nh = (int64_t*)malloc(sizeof(int64_t) * (size_t)(max_nh_batch + 1));
    int64_t y0[64];
    int status = 0;

    double mean = 0.0, var = 0.0, sum_sigf = 0.0;
    int64_t last_neval = 0, rmin = nmin, rmax = nmin, nbatch_count = 0;
    int64_t neval_batch = 0, hbase = 0;

    for (int64_t h = 0; h < nhcube; ++h) {
        int64_t ih = h - hbase;
        if (adaptive_sampler) {
            double v = sigf[h] * neval_sigf;
            if (v > 2147483647.0) v = 2147483647.0;
            int64_t n = (int64_t)(int)v + nmin;
            if (n > nmax) n = nmax;
            if (n < rmin) rmin = n;
            else if (n > rmax) rmax = n;
            nh[ih] = n;
            neval_batch += n;
        } else {
            nh[ih] = avg_nh;
            neval_batch += avg_nh;
        }
        if (neval_batch < P->min_neval_batch && h < nhcube - 1) continue;

        /* ---- build the batch ---- */
        int64_t p0 = last_neval;
        last_neval += neval_batch;
        int64_t nhb = h - hbase + 1;
        if (neval_batch > cap) {
            cap = 2 * neval_batch;
            y = (double*)realloc(y, sizeof(double) * (size_t)(cap * dim));
            x = (double*)realloc(x, sizeof(double) * (size_t)(cap * dim));
            U = (double*)realloc(U, sizeof(double) * (size_t)(cap * dim));
            jac = (double*)realloc(jac, sizeof(double) * (size_t)cap);
            fdv2 = (double*)realloc(fdv2, sizeof(double) * (size_t)cap);
            fx = (double*)realloc(fx, sizeof(double) * (size_t)cap);
        }
        if (P->rng_f32) vgo_uniforms_f32(P->seed, P->iteration, p0, neval_batch, dim, U);
        else vgo_uniforms_f64(P->seed, P->iteration, p0, neval_batch, dim, U);
        int64_t i0 = 0;
        for (int64_t k = 0; k < nhb; ++k) {
            int64_t t = hbase + k;
            for (int d = 0; d < dim; ++d) {
                y0[d] = t % P->nstrat[d];
                t = (t - y0[d]) / P->nstrat[d];
            }
            for (int d = 0; d < dim; ++d)
                for (int64_t i = i0; i < i0 + nh[k]; ++i)
                    y[i * dim + d] = ((double)y0[d] + U[i * dim + d]) / (double)P->nstrat[d];
            i0 += nh[k];
        }
        vgo_map(dim, P->ld, grid, inc, ninc, y, x, jac, neval_batch);
        i0 = 0;
        for (int64_t k = 0; k < nhb; ++k) {
            for (int64_t i = i0; i < i0 + nh[k]; ++i) jac[i] *= dv_y / (double)nh[k];
            i0 += nh[k];
        }
        /* ---- integrand (pyx:2126-2134) ---- */
        int rc = fcn(ctx, x, neval_batch, dim, fx);
        if (rc) { status = rc; goto done; }
        for (int64_t i = 0; i < neval_batch; ++i)
            if (isnan(fx[i])) { status = -3; goto done; }
        ++nbatch_count;

        /* ---- per-hcube two-pass sums (pyx:2138-2193) ---- */
        i0 = 0;
        for (int64_t k = 0; k < nhb; ++k) {
            int64_t n = nh[k];
            double sum_wf = 0.0, sum_dwf = 0.0, sum_dwf2 = 0.0;
            for (int64_t i = i0; i < i0 + n; ++i) sum_wf += jac[i] * fx[i];
            for (int64_t i = i0; i < i0 + n; ++i) {
                double dwf = jac[i] * fx[i] - sum_wf / (double)n;
                if (fabs(dwf) < eps * fabs(sum_wf / (double)n)) {
                    dwf = eps * fabs(sum_wf / (double)n);
                    sum_dwf2 += dwf * dwf;
                    dwf = 0.;
                } else {
                    sum_dwf2 += dwf * dwf;
                }
                sum_dwf += dwf;
                double t = jac[i] * fx[i] * (double)n;
                fdv2[i] = t * t;
            }
            mean += sum_wf + sum_dwf;
            var += ((double)n * sum_dwf2 - sum_dwf * sum_dwf) / ((double)n - 1.);
            double sigf2 = fabs(((double)n * sum_dwf2 - sum_dwf * sum_dwf) / ((double)n - 1.));
            if (adaptive_strat) {
                sigf[hbase + k] = pow(sigf2, P->beta / 2.);
                sum_sigf += sigf[hbase + k];
            }
            if (P->adapt_to_errors && P->adapt) {
                fdv2[i0 + n - 1] = sigf2;
                vgo_add_training_data(dim, P->ld, ninc, y + (i0 + n - 1) * dim,
                                      fdv2 + (i0 + n - 1), 1, sum_f, n_f);
            }
            i0 += n;
        }
        if (train_points) vgo_add_training_data(dim, P->ld, ninc, y, fdv2, neval_batch, sum_f, n_f);

        if (h < nhcube - 1) {
            neval_batch = 0;
            hbase = h + 1;
        }
    }
done:
    out->mean = mean;
    out->var = var;
    out->sum_sigf = sum_sigf;
    out->last_neval = last_neval;
    out->nh_min = adaptive_sampler ? rmin : nmin;
    out->nh_max = adaptive_sampler ? rmax : nmin;
    out->nbatch = nbatch_count;
    free(y); free(x); free(U); free(jac); free(fdv2); free(fx); free(nh);
    return status;
}

int vgo_run(const vgo_iter_params* P0, int nitn, double* grid, double* inc, int64_t* ninc,
            double* sigf, double* sum_sigf, vgo_fcn_t fcn, void* ctx, double* itn_mean,
            double* itn_var, int64_t* itn_neval) {
    vgo_iter_params P = *P0;
    const int dim = P.dim;
    const size_t nb = (size_t)dim * (size_t)(P.ld - 1);
    double* sum_f = (double*)malloc(sizeof(double) * nb);
    double* n_f = (double*)malloc(sizeof(double) * nb);
    double* new_grid = (double*)malloc(sizeof(double) * (size_t)dim * (size_t)P.ld);
    double* new_inc = (double*)malloc(sizeof(double) * nb);
    const int has_training = P.adapt && (P.adapt_to_errors || P.alpha > 0);
    int status = 0;
    for (int it = 0; it < nitn; ++it) {
        for (size_t i = 0; i < nb; ++i) { sum_f[i] = 0.0; n_f[i] = TINY; }
        P.iteration = P0->iteration + (uint32_t)it;
        vgo_iter_out o;
        status = vgo_iterate(&P, grid, inc, ninc, sigf, *sum_sigf, sum_f, n_f, fcn, ctx, &o);
        if (status) break;
        itn_mean[it] = o.mean;
        itn_var[it] = o.var;
        itn_neval[it] = o.last_neval;
        if (P.beta > 0 && !P.adapt_to_errors && P.adapt) {
            if (o.sum_sigf > 0) {
                *sum_sigf = o.sum_sigf;
            } else {
                for (int64_t h = 0; h < P.nhcube; ++h) sigf[h] = 1.;
                *sum_sigf = (double)(P.nhcube > 1 ? P.nhcube : 0);
            }
        }
        if (P.alpha > 0 && P.adapt) {
            vgo_adapt(dim, P.ld, grid, inc, ninc, has_training, sum_f, n_f, P.alpha, ninc,
                      P.ld, new_grid, new_inc);
            memcpy(grid, new_grid, sizeof(double) * (size_t)dim * (size_t)P.ld);
            memcpy(inc, new_inc, sizeof(double) * nb);
        }
    }
    free(sum_f); free(n_f); free(new_grid); free(new_inc);
    return status;
}

/* RAvg.add / _chi2 (pyx:2392-2410, 2344-2357).  The reference stores each
 * iteration as gvar(mean, sqrt(var)), so the variance it re-reads is
 * sqrt(var)**2. */
void vgo_ravg(int weighted, int n, const double* m, const double* v, double out[3]) {
    if (n <= 0) { out[0] = out[1] = out[2] = 0.0; return; }
    if (weighted) {
        double sw = 0.0, swm = 0.0;
        for (int k = 0; k < n; ++k) {
            double s = sqrt(v[k]);
            double vk = s * s;
            double w = 1. / (vk > TINY ? vk : TINY);
            sw += w;
            swm += w * m[k];
        }
        double var = 1. / sw;
        double mean = swm * var;
        double chi2 = 0.0;
        if (n > 1)
            for (int k = 0; k < n; ++k) {
                double s = sqrt(v[k]);
                double vk = s * s;
                double w = 1. / (vk > TINY ? vk : TINY);
                chi2 += (mean - m[k]) * (mean - m[k]) * w;
            }
        out[0] = mean; out[1] = sqrt(var); out[2] = chi2;
    } else {
        double ms = 0.0, vs = 0.0;
        for (int k = 0; k < n; ++k) {
            double s = sqrt(v[k]);
            ms += m[k];
            vs += s * s;
        }
        double mean = ms / n, var = vs / ((double)n * (double)n);
        double chi2 = 0.0;
        if (n > 1) {
            for (int k = 0; k < n; ++k) chi2 += (m[k] - mean) * (m[k] - mean);
            chi2 /= (vs / n);
        }
        out[0] = mean; out[1] = sqrt(var); out[2] = chi2;
    }
}
