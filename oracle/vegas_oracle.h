/*
 * vegas_oracle.h -- FP64 CPU oracle for the vegas+ hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (vegas_amd, libvgx.so) never
 * links, loads or calls it.
 *
 * Every routine restates the behaviour of gplepage/vegas 6.4.1
 * (src/vegas/_vegas.pyx, cited as pyx:N) in plain C99 double precision,
 * compiled with -O2 -ffp-contract=off so that the order and rounding of each
 * floating-point operation is the reference's (no fused multiply-add).
 *
 * Pinning status of each function (what each check fixes):
 *   vgo_philox4x32_10   Random123 known-answer vectors (SC'11 paper's KATs)
 *   vgo_uniforms_*      bit layout checked against independent numpy code
 *   vgo_make_uniform    numpy.linspace bit-for-bit
 *   vgo_map / invmap    tests/test_vegas.py:76-112 fixture, round trips
 *   vgo_add_training    brute-force numpy bincount restatement
 *   vgo_adapt           tests/test_vegas.py:39-63, 142-207 closed forms;
 *                       alpha > 0 smoothing and compression against the
 *                       printed grids of doc/source/eg2a.out (an ensemble
 *                       over seeded y; a dropped compression or an interior
 *                       smoothing weight of 2 instead of 6 fails it, a
 *                       weight of 4 does not: that weight is pinned only
 *                       against gross changes)
 *   vgo_geometry        tests/test_vegas.py:561-595, 703-772, eg10.out, eg6a.out
 *   vgo_iterate         constant-integrand known answers (tests:774-789,
 *                       1272-1296), analytic integrals, closed-form invariants
 *   RAvg (Python side)  tests/test_vegas.py:216-236 closed forms
 */
#ifndef VEGAS_ORACLE_H
#define VEGAS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (pyx:34-36) ---- */
double vgo_tiny(void);
double vgo_huge(void);
double vgo_epsilon(void);

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---- */
void vgo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Uniforms for points p0..p0+n-1 of one iteration, row-major U[(p-p0)*dim+d].
 * The counter layout is the one the GPU sampler uses:
 *   ctr = (p & 0xffffffff, p >> 32, iteration, block), key = (seed lo, seed hi)
 * f64: block j yields dims 2j, 2j+1, u = ((w[2e]>>5)*2^26 + (w[2e+1]>>6)) * 2^-53
 * f32: block j yields dims 4j..4j+3,  u = (w[e] >> 8) * 2^-24            */
void vgo_uniforms_f64(uint64_t seed, uint32_t iteration, int64_t p0, int64_t n,
                      int dim, double* U);
void vgo_uniforms_f32(uint64_t seed, uint32_t iteration, int64_t p0, int64_t n,
                      int dim, double* U);

/* ---- AdaptiveMap (pyx:39-599) ----
 * grid is dim x ld (ld >= max(ninc)+1), inc is dim x (ld-1), both row-major. */
void vgo_make_uniform(int dim, const double* lo, const double* hi,
                      const int64_t* ninc, int ld, double* grid, double* inc);
int vgo_map(int dim, int ld, const double* grid, const double* inc,
            const int64_t* ninc, const double* y, double* x, double* jac,
            int64_t n);
int vgo_invmap(int dim, int ld, const double* grid, const double* inc,
               const int64_t* ninc, const double* x, double* y, double* jac,
               int64_t n);
void vgo_jac1d(int dim, int ld, const double* inc, const int64_t* ninc,
               const double* y, double* jac, int64_t n);
/* sum_f/n_f are dim x (ld-1); caller initialises them to 0 / TINY (pyx:449-452) */
void vgo_add_training_data(int dim, int ld, const int64_t* ninc,
                           const double* y, const double* f, int64_t n,
                           double* sum_f, double* n_f);
/* adapt (pyx:467-594).  has_training=0 means sum_f is None.
 * new_grid is dim x new_ld with new_ld >= max(new_ninc)+1, new_inc dim x (new_ld-1).
 * Returns 0, or -1 for bad ninc. */
int vgo_adapt(int dim, int ld, const double* grid, const double* inc,
              const int64_t* ninc, int has_training, const double* sum_f,
              const double* n_f, double alpha, const int64_t* new_ninc,
              int new_ld, double* new_grid, double* new_inc);

/* ---- Integrator.set geometry (pyx:1331-1446) ---- */
typedef struct {
    int64_t nstrat[64];
    int64_t ninc[64];
    int64_t nhcube;
    int64_t neval;
    int64_t min_neval_hcube;
    int64_t workspace;
    int status; /* 0 ok, -1 ValueError, -2 MemoryError */
} vgo_geometry_out;
/* nstrat_in == NULL: automatic strata.  neval_given: 1 if neval was passed. */
void vgo_geometry(int dim, int64_t neval, int neval_given, const int64_t* nstrat_in,
                  double neval_frac, double beta, int adapt_to_errors,
                  int uniform_nstrat, int64_t maxinc_axis, double max_mem,
                  int minimize_mem, int64_t min_neval_batch,
                  int64_t max_neval_hcube, vgo_geometry_out* out);

/* ---- integrands ---- */
/* f(x) = sum_k A_k exp(-B_k sum_d (x_d - c_kd)^2) */
typedef struct {
    int dim;
    int ncomp;
    const double* A;
    const double* B;
    const double* c; /* ncomp x dim */
} vgo_gaussmix;
typedef int (*vgo_fcn_t)(void* ctx, const double* x, int64_t n, int dim, double* f);
int vgo_gaussmix_eval(void* ctx, const double* x, int64_t n, int dim, double* f);

/* ---- one vegas+ iteration (pyx:1650-1765 sampler, pyx:2096-2215 accumulation) ---- */
typedef struct {
    int dim;
    int ld;
    const int64_t* nstrat;
    int64_t nhcube;
    int64_t neval;
    double neval_frac;
    double beta;
    double alpha;
    int adapt;
    int adapt_to_errors;
    int64_t min_neval_hcube;
    int64_t max_neval_hcube;
    int64_t min_neval_batch;
    uint64_t seed;
    uint32_t iteration;
    int rng_f32;
} vgo_iter_params;

typedef struct {
    double mean;
    double var;
    double sum_sigf;     /* accumulated this iteration (pyx:2186) */
    int64_t last_neval;
    int64_t nh_min, nh_max;
    int64_t nbatch;
} vgo_iter_out;

/* sigf (nhcube) is read for allocation and updated in place when adaptive.
 * sum_f / n_f receive training data (caller initialises them: 0 / TINY).
 * sum_sigf_in is Integrator.sum_sigf from the previous iteration.
 * Returns 0, -3 if the integrand produced NaN (pyx:2133-2134), or the
 * integrand's own nonzero status. */
int vgo_iterate(const vgo_iter_params* P, const double* grid, const double* inc,
                const int64_t* ninc, double* sigf, double sum_sigf_in,
                double* sum_f, double* n_f, vgo_fcn_t fcn, void* ctx,
                vgo_iter_out* out);

/* Full Integrator.__call__ loop (pyx:2086-2227) for nitn iterations with
 * adapt and the sigf bookkeeping.  grid/inc/ninc/sigf updated in place; per-
 * iteration mean/var written to itn_mean/itn_var.  *sum_sigf in/out. */
int vgo_run(const vgo_iter_params* P, int nitn, double* grid, double* inc,
            int64_t* ninc, double* sigf, double* sum_sigf, vgo_fcn_t fcn,
            void* ctx, double* itn_mean, double* itn_var, int64_t* itn_neval);

/* RAvg weighted/unweighted average of the first n results (pyx:2392-2410,
 * 2344-2357).  out = {mean, sdev, chi2} */
void vgo_ravg(int weighted, int n, const double* m, const double* v, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
