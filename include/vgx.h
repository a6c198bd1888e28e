/*
 * vgx.h -- C-ABI of libvgx.so, the MI355X (gfx950) vegas+ hot path.
 *
 * Plain pointers and sizes only.  Every pointer argument documented as
 * "device" is device memory of the current HIP device; "host" pointers are
 * read synchronously before the call returns.  `stream` is a hipStream_t
 * (NULL = the null stream).  All kernels are enqueued asynchronously on that
 * stream; calls on one vgx_ctx are stream-ordered and not thread-safe.
 * No entry point returns memory the caller must free.  Every function returns
 * an int status: VGX_OK (0) or a negative VGX_E* code (see vgx_strerror).
 *
 * Layouts (row-major, as the reference's memoryviews, _vegas.pxd:21-31):
 *   grid   double[dim][ld]      nodes of axis d in grid[d][0..ninc[d]];
 *                               increments are grid[d][i+1]-grid[d][i]
 *                               (the reference stores them separately,
 *                               computed by exactly that subtraction)
 *   sum_f  double[dim][ld-1]    training sums   (AdaptiveMap.sum_f)
 *   n_f    double[dim][ld-1]    training counts (AdaptiveMap.n_f)
 *   y, x   T[n][dim]            points, T = double or float
 */
#ifndef VGX_H
#define VGX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VGX_MAXDIM 64
#define VGX_MAXCOMP 64  /* components of a multi-component integrand */

enum {
    VGX_OK = 0,
    VGX_EINVAL = -1,   /* bad argument (the reference's ValueError) */
    VGX_EHIP = -2,     /* HIP runtime error */
    VGX_ENAN = -3,     /* integrand evaluated to NaN (pyx:2133-2134) */
    VGX_ENOMEM = -4,   /* device allocation failed (MemoryError) */
    VGX_ERANGE = -5,   /* size beyond what the kernels support */
    VGX_ENODEV = -6,   /* no gfx950 device */
    VGX_ETIMEOUT = -7  /* a kernel's bounded wait gave up (never in a correct launch) */
};

const char* vgx_strerror(int status);
int vgx_version(void);

/* ---------------------------------------------------------------------- */
/* AdaptiveMap.map(y, x, jac, ny)              -- _vegas.pyx:310-360
 * ninc: host int32[dim].  grid: device.  y in, x/jac out: device.        */
int vgx_map_f64(int dim, const int32_t* ninc, const double* grid, int ld,
                const double* y, double* x, double* jac, int64_t n, void* stream);
/* Same map with float y/x/jac (index and interpolation computed in FP64). */
int vgx_map_f32(int dim, const int32_t* ninc, const double* grid, int ld,
                const float* y, float* x, float* jac, int64_t n, void* stream);

/* AdaptiveMap.invmap(x, y, jac, nx)           -- _vegas.pyx:362-416      */
int vgx_invmap_f64(int dim, const int32_t* ninc, const double* grid, int ld,
                   const double* x, double* y, double* jac, int64_t n, void* stream);

/* AdaptiveMap.jac1d(y)                        -- _vegas.pyx:265-295
 * jac out: device double[n][dim].                                        */
int vgx_jac1d_f64(int dim, const int32_t* ninc, const double* grid, int ld,
                  const double* y, double* jac, int64_t n, void* stream);

/* ---------------------------------------------------------------------- */
typedef struct vgx_ctx vgx_ctx;
/* Scratch owner for the current device (scan temporaries, tile partials,
 * staging).  Released by vgx_ctx_destroy.                                */
int vgx_ctx_create(vgx_ctx** out);
int vgx_ctx_destroy(vgx_ctx* ctx);

/* AdaptiveMap.add_training_data(y, f, ny)     -- _vegas.pyx:421-464
 * Adds |f[i]| and 1 to sum_f/n_f at bin floor(y[i][d]*ninc[d]) for every
 * axis d with 0 < y < 1.  ordered=1 sums every bin in point order, bit-for-
 * bit the reference's sequential loop; ordered=0 uses LDS-privatised
 * atomics (same counts, sums equal to rounding).                        */
int vgx_train_hist(vgx_ctx* ctx, int dim, const int32_t* ninc, int ld,
                   const double* y, const double* f, int64_t n,
                   double* sum_f, double* n_f, int ordered, void* stream);

/* No counterpart in the reference: vgx_train_hist's ordered mode for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments
 * and results (bit for bit) without a context: its temporaries live in a
 * static device array of the library, 128 MiB per GPU, held until the
 * process ends; VGX_ERANGE when the call needs more (about
 * dim*(8 + 4*ld/1024) bytes per point plus 8*dim*ld).  Every call on one
 * GPU shares that array, so calls that overlap on one GPU race.  Allocates
 * nothing and launches on `stream`.                                        */
int vgx_train_hist_noalloc(int dim, const int32_t* ninc, int ld, const double* y, const double* f,
                           int64_t n, double* sum_f, double* n_f, void* stream);

/* AdaptiveMap.adapt(alpha, ninc)              -- _vegas.pyx:467-594
 * has_training=0 is the reference's sum_f None.  new_grid: device
 * double[dim][new_ld], new_ld >= max(new_ninc)+1.  exact=1 runs the
 * reference's sequential sums and CDF walk in one lane per axis (order of
 * operations identical to the reference); exact=0 uses block-parallel
 * reductions and a parallel CDF inversion (agrees to ~1e-13 relative).  */
int vgx_adapt(vgx_ctx* ctx, int dim, const int32_t* ninc, const double* grid, int ld,
              int has_training, const double* sum_f, const double* n_f, double alpha,
              const int32_t* new_ninc, double* new_grid, int new_ld, int exact,
              void* stream);

/* ---------------------------------------------------------------------- */
/* Integrator: one vegas+ run of nitn iterations enqueued without host
 * synchronisation.  Geometry (pyx:1331-1446) is computed by the caller. */
enum { VGX_PREC_F64 = 0, VGX_PREC_F32 = 1 };
enum { VGX_FCN_GAUSSMIX = 1, VGX_FCN_DIAGRIDGE = 2, VGX_FCN_BUFFER = 3 };
enum {
    VGX_FLAG_EXACT_ADAPT = 1,  /* vgx_adapt(exact=1) inside the loop, and the training
                                  histogram summed in point order (pyx:457-463) */
    VGX_FLAG_EXACT_DIV = 2,    /* y = (c+u)/s by IEEE division (as pyx:1741) */
    VGX_FLAG_UNORDERED_HIST = 4 /* training histogram by LDS atomics in timing order:
                                  faster, but not reproducible bit for bit */
};

typedef struct {
    int32_t dim;
    int32_t ld;                      /* grid row stride, >= max(ninc)+1 */
    int32_t ninc[VGX_MAXDIM];
    int32_t nstrat[VGX_MAXDIM];
    int64_t nhcube;
    int64_t neval;
    double neval_frac;               /* raw Integrator.neval_frac */
    double alpha;
    double beta;
    int32_t adapt;
    int32_t adapt_to_errors;
    int64_t min_neval_hcube;
    int64_t max_neval_hcube;
    uint64_t seed;
    uint32_t iteration;              /* RNG stream of the first iteration */
    int32_t nitn;
    int32_t precision;               /* VGX_PREC_* : integrand arithmetic */
    int32_t flags;                   /* VGX_FLAG_* */
    int32_t weighted;                /* RAvg weighted average (adapt) */
    int32_t reserved;
    /* Caller-supplied uniforms (Integrator.ran_array_generator, pyx:1676-1680,
     * 1732): device double[total][dim], row p holding the dim numbers in [0,1)
     * of point p of the iteration (p as vgx_iter_begin orders the points), so
     * y = (stratum + u)/nstrat (pyx:1741) in FP64.  Staged iterations only
     * (vgx_iter_sample/accumulate/end; precision VGX_PREC_F64, nitn = 1);
     * vgx_integrate rejects it.  NULL: the library's Philox4x32-10 stream.  */
    const double* uniforms;
} vgx_config;

/* Device integrands.  GAUSSMIX: f(x) = sum_k A[k] exp(-B[k] |x - c[k]|^2)
 * with c double[ncomp][dim].  DIAGRIDGE: the same mixture with centres on
 * the main diagonal, c[k] = t[k]*(1,...,1) (c holds t[ncomp]); evaluated
 * with |x-c|^2 = P + dim (xbar - t)^2, xbar the mean coordinate and
 * P = sum_d (x_d - xbar)^2 computed once per point.  BUFFER: f values supplied
 * by the caller (device double, indexed by global point index - fbase).  */
typedef struct {
    int32_t kind;
    int32_t ncomp;
    const double* A;                 /* device */
    const double* B;                 /* device */
    const double* c;                 /* device */
    const double* fbuf;              /* BUFFER only */
    int64_t fbase;
} vgx_integrand;

/* Caller-owned device state, updated in place. */
typedef struct {
    double* grid;                    /* dim x ld */
    double* sigf;                    /* nhcube (stratification weights) */
    double* sum_sigf;                /* [1] Integrator.sum_sigf */
    double* sum_f;                   /* dim x (ld-1) training scratch */
    double* n_f;                     /* dim x (ld-1) */
    double* itn_mean;                /* [nitn] per-iteration estimate */
    double* itn_var;                 /* [nitn] */
    double* ravg;                    /* [4] mean, sdev, chi2, sum_w (FP64) */
    int64_t* itn_neval;              /* [nitn] last_neval */
    int64_t* itn_nhrange;            /* [2*nitn] neval_hcube_range */
    int32_t* status;                 /* [1] 0 or VGX_ENAN */
} vgx_state;

/* Integrator.__call__ for a device integrand   -- _vegas.pyx:2086-2227.
 * Runs cfg->nitn iterations: allocation (pyx:1657-1706), stratified y,
 * map, integrand, per-hcube two-pass sums with the EPSILON guard
 * (pyx:2138-2193), training histogram, sigf/sum_sigf update, adapt and the
 * running average, all on device.  Host synchronisation is the caller's. */
int vgx_integrate(vgx_ctx* ctx, const vgx_config* cfg, const vgx_integrand* fcn,
                  vgx_state* st, void* stream);

/* Staged iteration for host-driven (Python) integrands, in three steps:
 * vgx_iter_begin    allocation + scan; writes n_h (device int32[nhcube]) so
 *                   the caller can cut batches at hcube boundaries exactly
 *                   as _random_batch does (pyx:1692-1710);
 * vgx_iter_sample   x (and optionally y, wgt, hcube) for the hcubes
 *                   [h0, h1) (pyx:1731-1759);
 * vgx_iter_accumulate  per-hcube sums and training data for the same batch
 *                   given f (device double[n]) (pyx:2136-2197);
 * vgx_iter_end      closes the iteration: estimate, sum_sigf, running
 *                   average and adapt (pyx:2199-2217).  it = iteration
 *                   index within cfg->nitn.                              */
int vgx_iter_begin(vgx_ctx* ctx, const vgx_config* cfg, vgx_state* st, int it,
                   int32_t* n_h, int64_t* total_host, void* stream);
int vgx_iter_sample(vgx_ctx* ctx, const vgx_config* cfg, vgx_state* st, int it,
                    int64_t h0, int64_t h1, int64_t p0, int64_t p1,
                    double* x, double* y, double* wgt, int64_t* hcube, void* stream);
int vgx_iter_accumulate(vgx_ctx* ctx, const vgx_config* cfg, vgx_state* st, int it,
                        int64_t h0, int64_t h1, int64_t p0, int64_t p1,
                        const double* f, void* stream);
int vgx_iter_end(vgx_ctx* ctx, const vgx_config* cfg, vgx_state* st, int it,
                 void* stream);
/* Multi-component integrands (fcn_size m > 1; pyx:2136-2183 and its
 * correlate_integrals switch): for the batch [h0,h1) / [p0,p1) of the
 * iteration opened by vgx_iter_begin, f is device double[n][m] (n = p1-p0)
 * and wgt the weights vgx_iter_sample gave for the same points.  Adds the
 * batch's sums to acc (device double, double-double (hi, lo) pairs, zeroed
 * by the caller per iteration): mean[m], then the covariance lower triangle
 * row by row ((0,0) (1,0) (1,1) (2,0) ...: m(m+1)/2 pairs) when correlate,
 * else the m variances.  Component 0 alone steers the sampling: the caller
 * passes f[:,0] to vgx_iter_accumulate for sigf and the training data.
 * status (device int32, optional): VGX_ERANGE when an hcube's points are
 * not inside [p0, p1) (those hcubes are skipped, nothing is read).
 * Replaces the fcn_size loops of Integrator.__call__ (pyx:2146-2183).   */
int vgx_iter_accumulate_cov(vgx_ctx* ctx, const vgx_config* cfg, int64_t h0, int64_t h1,
                            int64_t p0, int64_t p1, const double* f, const double* wgt,
                            int m, int correlate, double* acc, int32_t* status, void* stream);

/* Kernel timing (measurement; no counterpart in the reference).
 * vgx_set_timing(ctx, 1) clears the record and brackets the following
 * launches (up to 256 in all) of the timed kernels with HIP events on their
 * stream: the fused sampler (tag 1), the batched sampler (2; while timing
 * is on, vgx_batch_integrate launches its kernels directly instead of
 * through its graph), the flow's backward with its weight gradients (3) and
 * sampling (6) kernels, and the training-histogram rows (5).
 * vgx_timed_launches writes the durations in ms and the tags in launch
 * order and returns how many were recorded; vgx_fused_times writes those
 * of the fused sampler only.  Both synchronise on the events they read. */
int vgx_set_timing(vgx_ctx* ctx, int on);
int vgx_timed_launches(vgx_ctx* ctx, float* ms, int32_t* tags, int cap);
int vgx_fused_times(vgx_ctx* ctx, float* ms, int cap);

/* ---------------------------------------------------------------------- */
/* Batched integrals: n independent vegas+ integrations of single-Gaussian
 * integrands f_i(x) = A_i exp(-B_i |x - c_i|^2) on [0,1]^dim_i, advanced
 * together: one allocation, one persistent sampler launch (work items of
 * whole wave tiles, each inside one integral), one straddler pass, one
 * close and one adapt per iteration for all of them.  The nitn iterations
 * of a call are captured once as a hipGraph and replayed by later calls
 * with the same arguments.  No counterpart in the reference, which runs one
 * Integrator at a time; each integral follows Integrator.__call__
 * (pyx:1912-2230) with adapt_to_errors=False and beta > 0, and its samples
 * are the single-integral Philox stream (same seed, iteration and point
 * index), so the batch reproduces n separate vgx_integrate runs.          */
#define VGX_BATCH_MAXDIM 10
#define VGX_BATCH_MAXINC 500
#define VGX_BATCH_LD (VGX_BATCH_MAXINC + 1)

typedef struct vgx_batch vgx_batch;

typedef struct {
    int32_t n;                        /* integrals */
    const int32_t* dim;               /* host [n], 1..VGX_BATCH_MAXDIM */
    const int32_t* ninc;              /* host [n][VGX_BATCH_MAXDIM], <= MAXINC */
    const int32_t* nstrat;            /* host [n][VGX_BATCH_MAXDIM] */
    const int64_t* neval;             /* host [n] */
    const int64_t* min_neval_hcube;   /* host [n] */
    const int64_t* max_neval_hcube;   /* host [n] */
    const uint64_t* seed;             /* host [n] */
    double neval_frac;
    double alpha;
    double beta;                      /* > 0 */
    int32_t tiles_per_item;           /* work-item size in wave tiles (1..240; 0 = 128) */
    int32_t blocks;                   /* sampler workgroups (0 = one per CU) */
} vgx_batch_config;

/* Caller-owned device state (sizes from vgx_batch_sizes). */
typedef struct {
    double* grid;        /* [n][VGX_BATCH_MAXDIM][VGX_BATCH_LD] */
    double* sigf;        /* [total]: integral i's hcubes from hcube_base[i],
                            one spare entry after each integral */
    double* sum_sigf;    /* [n] */
    double* itn_mean;    /* [nitn][n] */
    double* itn_var;     /* [nitn][n] */
    int64_t* itn_neval;  /* [nitn][n] */
    double* ravg;        /* [n][4]: mean, sdev, chi2, dof of the call */
    int32_t* status;     /* [1]: 0, VGX_ENAN or VGX_ERANGE */
} vgx_batch_state;

int vgx_batch_create(vgx_ctx* ctx, const vgx_batch_config* cfg, vgx_batch** out);
int vgx_batch_destroy(vgx_batch* b);
/* Length of sigf (sum of nhcube_i + 1), and hcube_base[n] (host, may be NULL). */
int vgx_batch_sizes(const vgx_batch* b, int64_t* total_hcubes, int64_t* hcube_base);
/* Integrand parameters: A, B device double[n]; c device double[n][MAXDIM]. */
int vgx_batch_integrate(vgx_batch* b, const double* A, const double* B, const double* c,
                        const vgx_batch_state* st, int nitn, void* stream);
/* The next call's first iteration number (Philox stream): calls continue it. */
int vgx_batch_set_iteration(vgx_batch* b, uint32_t iteration);

/* ---------------------------------------------------------------------- */
/* Learned map: an affine-coupling normalizing flow on [0,1]^dim (dim <= 15,
 * nlayers <= 8), z -> u0 = logit(z) -> couplings -> x = sigmoid(u).  Layer
 * k conditions on the dims in amask[k]; its conditioner is the MLP
 * (dim -> 128 -> 128 -> 2 dim) with bf16 weights and FP32 accumulation on
 * MFMA; the log-determinant is FP32.  No counterpart in the reference, whose
 * only map is the separable grid (pyx:39-832); the flow stands in for it on
 * non-separable integrands, and is used only when its measured per-sample
 * variance is lower.  Parameters (FP32 master, per layer:
 * W1[128][16] (column 15 is the bias) | W2[128][128] | b2[128] | W3[32][128]
 * | b3[32]; rows 0..15 of W3 give s, rows 16..31 give t).
 * Points are float [n][16] (dims padded to 16).                           */
typedef struct vgx_flow vgx_flow;
int vgx_flow_create(int dim, int nlayers, const uint32_t* amask, vgx_flow** out);
int vgx_flow_destroy(vgx_flow* fl);
int vgx_flow_nparams(const vgx_flow* fl, int64_t* n);
/* host float[nparams]; resets the Adam moments */
int vgx_flow_set_params(vgx_flow* fl, const float* host, void* stream);
/* host copies of the parameters and of the last gradient (either may be NULL) */
int vgx_flow_get_params(const vgx_flow* fl, float* host, float* grad_host, void* stream);
/* Importance-sampling estimate: n points z (Philox: key seed, counter
 * (base + i, iteration, lane group)), x = F(z), f(x) J(z) summed:
 * sums[0] = sum f J, sums[1] = sum (f J)^2 (host; the call then waits for
 * the stream).  sums may be NULL when xout and fout are given: nothing is
 * read back and the call returns without waiting.  fcn: GAUSSMIX or
 * DIAGRIDGE with device FP64 parameters (evaluated in FP32).  xout, fout
 * (device, optional): the points and f J.  status: device int32, set to
 * VGX_ENAN when f is NaN.                                                 */
int vgx_flow_sample(vgx_flow* fl, const vgx_integrand* fcn, int64_t n, uint64_t base, uint64_t seed,
                    uint32_t iteration, double* sums, float* xout, float* fout, int32_t* status,
                    void* stream);
/* log q(x) (device float[n]) of points x (device float[n][16]); u0
 * (device float[n][16], optional): logit(z) of their preimages.           */
int vgx_flow_logq(vgx_flow* fl, const float* x, int64_t n, float* logq, float* u0, void* stream);
/* Gradient of  L = -sum_i c_i log q(x_i)  for samples x (device
 * float[n][16]) with integrand values f and log sampling weights log w
 * (device float[n]), c normalised to sum 1:
 *   VGX_FLOW_LOSS_VARIANCE  c_i ~ w_i f_i^2 / q(x_i): the gradient of
 *                           sum_i w_i f_i^2 / q(x_i) = Var(f/q) + I^2 (est.),
 *                           up to a positive factor;
 *   VGX_FLOW_LOSS_KL        c_i ~ w_i |f_i|: forward KL from |f|/I to q;
 *   VGX_FLOW_LOSS_SQUARED   c_i ~ w_i f_i^2: the variance loss for samples of
 *                           the flow itself, given f J (vgx_flow_sample's
 *                           fout) as f and log w = 0 (then w f^2 / q is
 *                           (f J)^2 / n: no log q is needed; likewise the KL
 *                           loss takes f J and log w = 0 for them).
 * Kept on the device for vgx_flow_adam (vgx_flow_get_params reads it); the
 * per-sample terms are summed in a fixed order, so the same inputs give the
 * same bits.                                                              */
enum { VGX_FLOW_LOSS_VARIANCE = 0, VGX_FLOW_LOSS_KL = 1, VGX_FLOW_LOSS_SQUARED = 2 };
int vgx_flow_grad(vgx_flow* fl, const float* x, const float* f, const float* logw, int64_t n, int loss,
                  void* stream);
/* One Adam step with the last gradient; bf16 copies refreshed.            */
int vgx_flow_adam(vgx_flow* fl, float lr, float beta1, float beta2, float eps, float clip, void* stream);
/* f at points x (device float[n][16]) into f (device float[n]).           */
int vgx_flow_eval(vgx_flow* fl, const vgx_integrand* fcn, const float* x, int64_t n, float* f,
                  void* stream);
/* D = A B for 16x16 bf16 A, B (device, row-major) on one MFMA: checks the
 * instruction's lane layout.                                              */
int vgx_flow_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, void* stream);

#ifdef __cplusplus
}
#endif
#endif
