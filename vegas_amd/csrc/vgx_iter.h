// vgx_iter.h -- shared definitions of the vegas+ iteration kernels.
#pragma once
#include "vgx_internal.h"

namespace vgx {

constexpr int kK = 4;               // points per lane in a wave tile
constexpr int kT = 64 * kK;         // points per wave tile
constexpr int kNoff = kT / 2 + 4;   // offsets slice: hcubes touching a tile (+1 end, +pad)
constexpr int kFusedThreads = 1024; // 16 waves per workgroup, one workgroup per CU
constexpr int kAllocThreads = 256;
constexpr int kAllocItems = 8;
constexpr int kAllocPerBlock = kAllocThreads * kAllocItems;
constexpr uint32_t kCntFlushTiles = 224; // u16 LDS counts: (224 + 15) tiles x 256 < 65536
constexpr int kStradThreads = 256;  // straddler kernel: one lane per tile slot
constexpr int kLaneParts = 8;       // cut hcubes of up to this many tiles: one lane; more: the wave
constexpr int kKeySentinel = 0x7fff;

// Register-array bound of a dimension template argument: DIM > 0 is exact;
// DIM <= 0 is a runtime dimension of at most -DIM (all 64 for DIM = 0).
__host__ __device__ constexpr int nd_max(int DIM) { return DIM > 0 ? DIM : (DIM < 0 ? -DIM : VGX_MAXDIM); }

// Kernel-argument copy of the iteration geometry (Integrator.set, pyx:1331-1446).
struct IterGeom {
    int dim, ld, nb;
    int adaptive_sampler;  // beta > 0 && nhcube > 1 && !adapt_to_errors  (pyx:1675)
    int adaptive_strat;    // ... && adapt                              (pyx:2052-2055)
    int train_points;      // !adapt_to_errors && adapt && alpha > 0      (pyx:2196)
    int train_errors;      // adapt_to_errors && adapt                    (pyx:2187)
    int rng32;             // 24-bit uniforms (FP32 integrand) else 53-bit
    int exact_div;         // y = (c+u)/s by IEEE division
    int hist_lds, grid_lds;
    int hist_unordered;    // hist_lds: LDS atomics in timing order (VGX_FLAG_UNORDERED_HIST)
    int weighted;
    int adapt;
    int fast_pow;          // sigf = |var|^(beta/2) in the sampler, FP32 log2/exp2
                           // (else FP64 pow in a separate pass)
    int per_all;           // rng32, no exact division, and ninc a multiple of
                           // nstrat on every axis: the integer stratum->bin path
    int32_t ninc[VGX_MAXDIM];
    int32_t nstrat[VGX_MAXDIM];
    double inv_nstrat[VGX_MAXDIM];
    DivMagic dm[VGX_MAXDIM];
    int64_t nhcube;
    double neval;
    double neval_frac;
    double beta_half;
    double beta;
    double dv_y;
    double jn;             // product over the axes of (double)ninc
    int32_t nmin, nmax, avg_nh, pad2;
    uint32_t seed_lo, seed_hi;
    uint32_t iteration;  // RNG stream of this iteration
    uint32_t itn;        // index within the run
    uint32_t nitn;
    uint32_t pad3;
    const double* uext;  // caller's uniforms [total][dim] (53-bit path), or null: Philox
};

// Device-resident scalars of one run.
struct Scalars {
    uint32_t total;       // points this iteration (off[nhcube])
    uint32_t overflow;
    int32_t nh_min, nh_max;
    // current batch (whole iteration for device integrands)
    uint32_t b_p0, b_p1, b_h0, b_h1, b_ntiles, pad0;
    uint64_t wfn_max_bits;  // max |w f n| seen (non-negative double bits)
    double acc[6];          // iteration sums (DD): mean, var, sum_sigf
    int32_t status;         // 0 or VGX_ENAN
    int32_t reset_sigf;
    double pad1[2];
};

// The part of an hcube cut by a tile edge that lies in one tile: its point
// count and its two-pass sums about the part's own mean S/n (the straddler
// kernel combines the parts exactly, Chan-Golub-LeVeque style).
struct Slot {
    int32_t h;      // hcube, or -1
    uint32_t n;     // points of h in this tile
    double S;       // sum of w f over them
    double Q;       // sum of guarded (w f - S/n)^2
    double D;       // sum of guarded (w f - S/n)
};

struct AxisC;
struct FusedPtrs {
    const AxisC* axes;  // dim per-axis sampler constants (device)
    const double* grid;
    const uint32_t* off;
    const int32_t* tile_h;
    double* sigf;
    Slot* slots;        // [2 * tile]: part of the hcube cut by the tile's start / end
    double* tile_part;  // [tile][3]: the tile's sums (mean, variance, sigf), summed in tile order
    double* blk_part;   // [gridDim.x][6]
    double* sum_f;
    double* n_f;
    double* hpart_s;    // LDS histogram (fast mode): [gridDim.x][dim * nb] sums and counts
    uint32_t* hpart_c;  //   of each workgroup, summed in workgroup order after the launch
    uint32_t* rec_bins; // training records, one per point of the batch (index i = point
    double* rec_c2;     //   index - b_p0): the bins of axes 2k, 2k+1 as the u16 halves of
                        //   rec_bins[k * rec_stride + i] (kNoBin where the point is not
                        //   trained on the axis, pyx:459-463), (w f n)^2 (pyx:2173) at rec_c2[i]
    uint32_t rec_stride; // a multiple of the tile size (whole tiles per row)
    double* errv;       // adapt_to_errors: |variance| per hcube
    Scalars* sc;
};

// ---------------------------------------------------------------------------
// Integrands (device functors).  eval() receives x in FP64 and returns f in
// FP64; the arithmetic precision is the functor's.

// Reads a parameter that no kernel writes while the integrand runs, through
// the constant address space: at a wave-uniform index (the component loop)
// this is a scalar load into SGPRs rather than a vector load with its full
// memory latency per component.
template <class T>
__device__ __forceinline__ T ldc(const T* p, int i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// f(x) = sum_k A_k exp(-B_k |x - c_k|^2)
struct GaussMix64 {
    static constexpr bool kRng32 = false;
    int ncomp;
    const double* A;
    const double* B;
    const double* c;
    template <int DIM>
    __device__ double eval(const double* x, int dim, uint32_t) const {
        const int nd = DIM > 0 ? DIM : dim;
        double acc = 0.0;
        for (int k = 0; k < ncomp; ++k) {
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < nd_max(DIM); ++d) {
                if (DIM <= 0 && d >= nd) break;
                double t = x[d] - ldc(c, k * nd + d);
                r2 += t * t;
            }
            acc += ldc(A, k) * exp(-ldc(B, k) * r2);
        }
        return acc;
    }
};

// A * 2^y for y <= 0 in FP32 arithmetic with the FP64 exponent range: y = n + f
// (n = floor(y), f in [0,1), an exact split), A * 2^f by the FP32 hardware
// exp2, and the scaling by 2^n exact in FP64.  FP32 exp flushes e^-104 and
// below to 0; here a point far from a narrow peak keeps its tiny value (down
// to 2^-1074) as the reference's FP64 evaluation does -- the training signal
// that leads the map to the peak in the first iterations.  NaN propagates.
__device__ __forceinline__ double scaled_exp2(float A, float y) {
    y = y < -1100.0f ? -1100.0f : y;  // 2^-1100 is 0 in FP64 as well
    const float fl = floorf(y);
    return ldexp((double)(A * __builtin_amdgcn_exp2f(y - fl)), (int)fl);
}

struct GaussMix32 {
    static constexpr bool kRng32 = true;
    int ncomp;
    const float* A;
    const float* B;
    const float* c;
    template <int DIM>
    __device__ double eval(const double* x, int dim, uint32_t) const {
        const int nd = DIM > 0 ? DIM : dim;
        float xf[nd_max(DIM)];
#pragma unroll
        for (int d = 0; d < nd_max(DIM); ++d) {
            if (DIM <= 0 && d >= nd) break;
            xf[d] = (float)x[d];
        }
        // the components' values, each with the FP64 exponent range, summed in FP64
        double acc = 0.0;
        for (int k = 0; k < ncomp; ++k) {
            float r2 = 0.0f;
#pragma unroll
            for (int d = 0; d < nd_max(DIM); ++d) {
                if (DIM <= 0 && d >= nd) break;
                float t = xf[d] - ldc(c, k * nd + d);
                r2 = __builtin_fmaf(t, t, r2);
            }
            acc += scaled_exp2(ldc(A, k), (-ldc(B, k) * r2) * 1.44269504088896341f);
        }
        return acc;
    }
};

// Mixture with centres t_k (1,...,1): |x - c_k|^2 = P + dim (xbar - t_k)^2,
// P = sum_d (x_d - xbar)^2 (exact algebra, no cancellation).
struct DiagRidge32 {
    static constexpr bool kRng32 = true;
    int ncomp;
    const float* A;
    const float* B;
    const float* c;  // t_k
    template <int DIM>
    __device__ double eval(const double* x, int dim, uint32_t) const {
        const int nd = DIM > 0 ? DIM : dim;
        float xf[nd_max(DIM)];
        float s = 0.0f;
#pragma unroll
        for (int d = 0; d < nd_max(DIM); ++d) {
            if (DIM <= 0 && d >= nd) break;
            xf[d] = (float)x[d];
            s += xf[d];
        }
        const float xbar = s / (float)nd;
        float P = 0.0f;
#pragma unroll
        for (int d = 0; d < nd_max(DIM); ++d) {
            if (DIM <= 0 && d >= nd) break;
            float t = xf[d] - xbar;
            P = __builtin_fmaf(t, t, P);
        }
        const float fd = (float)nd;
        float acc = 0.0f;
        // eight components per step with their constants loaded first (wide
        // scalar loads, one wait instead of two per component); the sum runs
        // in component order as in the remainder loop
        constexpr int kStep = 8;
        int k = 0;
        for (; k + kStep <= ncomp; k += kStep) {
            float cc[kStep], aa[kStep], bb[kStep];
#pragma unroll
            for (int i = 0; i < kStep; ++i) {
                cc[i] = ldc(c, k + i);
                aa[i] = ldc(A, k + i);
                bb[i] = ldc(B, k + i);
            }
#pragma unroll
            for (int i = 0; i < kStep; ++i) {
                float t = xbar - cc[i];
                float r2 = __builtin_fmaf(fd * t, t, P);
                acc = __builtin_fmaf(aa[i], __expf(-bb[i] * r2), acc);
            }
        }
        for (; k < ncomp; ++k) {
            float t = xbar - ldc(c, k);
            float r2 = __builtin_fmaf(fd * t, t, P);
            acc = __builtin_fmaf(ldc(A, k), __expf(-ldc(B, k) * r2), acc);
        }
        return (double)acc;
    }
};

struct DiagRidge64 {
    static constexpr bool kRng32 = false;
    int ncomp;
    const double* A;
    const double* B;
    const double* c;
    template <int DIM>
    __device__ double eval(const double* x, int dim, uint32_t) const {
        const int nd = DIM > 0 ? DIM : dim;
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < nd_max(DIM); ++d) {
            if (DIM <= 0 && d >= nd) break;
            s += x[d];
        }
        const double xbar = s / (double)nd;
        double P = 0.0;
#pragma unroll
        for (int d = 0; d < nd_max(DIM); ++d) {
            if (DIM <= 0 && d >= nd) break;
            double t = x[d] - xbar;
            P += t * t;
        }
        double acc = 0.0;
        for (int k = 0; k < ncomp; ++k) {
            double t = xbar - ldc(c, k);
            acc += ldc(A, k) * exp(-ldc(B, k) * (P + (double)nd * t * t));
        }
        return acc;
    }
};

// f supplied by the caller for global point p (host-driven integrands).
template <bool R32>
struct BufferFcn {
    static constexpr bool kRng32 = R32;
    const double* f;
    int64_t base;
    template <int DIM>
    __device__ double eval(const double*, int, uint32_t p) const {
        return f[(int64_t)p - base];
    }
};

// AdaptiveMap.adapt for the integrator loop (vgx_adapt.hip): in place, a
// no-op once *status != 0, clears sum_f/n_f after reading them.
int adapt_in_place(vgx_ctx* ctx, int dim, const int32_t* ninc, double* grid, int ld, double* sum_f,
                   double* n_f, int has_training, double alpha, int exact, const int32_t* status,
                   hipStream_t st);
// Its set-up alone: fills *args and sets *parallel when the workgroup-per-
// axis kernel applies (the caller then runs adapt_parallel_axis itself).
struct AdaptArgs;
int adapt_in_place_setup(vgx_ctx* ctx, int dim, const int32_t* ninc, double* grid, int ld, double* sum_f,
                         double* n_f, int has_training, double alpha, int exact, const int32_t* status,
                         AdaptArgs* args, int* parallel);

// Launchers, one per (integrand, dim) instantiation (vgx_fused_*.hip).
// ev0/ev1 (may be null): events that take the launch's own start and end
// timestamps (hipExtLaunchKernel), so timing adds no packets to the stream.
template <class Fcn>
int launch_fused(const IterGeom& g, const FusedPtrs& io, const Fcn& fcn, int grid_blocks,
                 size_t lds, hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

}  // namespace vgx
