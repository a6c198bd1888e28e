// vgx_flow.hip -- the learned map: an affine-coupling normalizing flow on
// [0,1]^D (D <= 15) with hand-written BF16 MFMA kernels (gfx950
// v_mfma_f32_16x16x16_bf16, FP32 accumulation, FP32 log-det).
//
// Map (the same arithmetic as oracle/flow_oracle.py): z -> u0 = logit(z) ->
// L couplings -> x = sigmoid(uL).  Coupling k conditions on the set A_k:
//   a = (u_d, d in A_k; 0 else; a_15 = 1),  h1 = relu(W1 a),
//   h2 = relu(W2 h1 + b2),  o = W3 h2 + b3,  s = 4 tanh(o_0..15 / 4),  t = o_16..31
//   u_B <- u_B e^s + t (forward),  u_B <- (u_B - t) e^-s (inverse).
//
// MFMA orientation.  Every GEMM is computed transposed, C[out][sample] =
// W[out][in] X[in][sample], on tiles of 16 samples.  With the 16x16x16
// instruction, lane l holds C[(l/16)*4 + i][l%16] (i < 4) and the B operand
// X[(l/16)*4 + j][l%16]: one layer's output tile mt, after bias, ReLU and the
// bf16 rounding, is the next layer's k-step mt in registers, with no data
// movement.  The coupling state u (dims (l/16)*4 + i of sample l%16) sits in
// the same lanes as s (output tile 0) and t (output tile 1), so the update is
// lane-local; per-sample sums reduce across the four lane groups.
//
// Kernels:
//   flow_sample   z (Philox) -> x, log J, f(x), sums of f J and (f J)^2: the
//                 estimator; layers streamed through LDS, samples in registers
//   flow_inverse  x -> z, log q(x), each layer's input saved (training)
//   flow_lossw    variance-loss weights c_i ~ w_i f_i^2 / q(x_i), normalised
//   flow_backward one layer: recompute, backprop, and the weight gradients
//                 (MFMA over the sample axis, operands staged in LDS)
//   flow_wgrad_reduce  the workgroups' weight-gradient sums, in a fixed order
//   flow_adam     Adam on the FP32 master weights, bf16 copies refreshed
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "vgx_internal.h"

namespace vgx {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kFH = 128;             // hidden width
constexpr int kFI = 16;              // padded input width, column 15 = layer-1 bias
constexpr int kFO = 32;              // output rows: s 0..15, t 16..31
constexpr int kFMaxL = 8;            // layers
constexpr float kSMax = 4.0f;
// FP32 master parameters of one layer: W1 | W2 | b2 | W3 | b3
constexpr int kPW1 = 0, kPW2 = kPW1 + kFH * kFI, kPB2 = kPW2 + kFH * kFH, kPW3 = kPB2 + kFH,
              kPB3 = kPW3 + kFO * kFH, kPLayer = kPB3 + kFO;
// bf16 weights of one layer: W1 | W2 | W3, and the transposes W1t | W2t | W3t,
// with padded row strides so that the 16 rows of an MFMA A fragment fall in
// distinct LDS banks (a 256-byte stride puts them all in one bank)
constexpr int kS1 = 20, kS2 = 132, kS3 = 132;          // W1 [128][16], W2 [128][128], W3 [32][128]
constexpr int kT1 = 132, kT2 = 132, kT3 = 36;          // W1t [16][128], W2t [128][128], W3t [128][32]
constexpr int kHW1 = 0, kHW2 = kHW1 + kFH * kS1, kHW3 = kHW2 + kFH * kS2, kHLayer = kHW3 + kFO * kS3;
constexpr int kTW1 = 0, kTW2 = kTW1 + kFI * kT1, kTW3 = kTW2 + kFH * kT2;
static_assert(kTW3 + kFH * kT3 <= kHLayer, "transposed layout fits the layer block");

// FP32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ inline bf16x4_t pack_bf(const f32x4_t v) {
    bf16x4_t r;
    r[0] = __builtin_bit_cast(short, (__bf16)v[0]);
    r[1] = __builtin_bit_cast(short, (__bf16)v[1]);
    r[2] = __builtin_bit_cast(short, (__bf16)v[2]);
    r[3] = __builtin_bit_cast(short, (__bf16)v[3]);
    return r;
}

__device__ inline f32x4_t mfma16(bf16x4_t a, bf16x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two k-steps in one gfx950 v_mfma_f32_16x16x32_bf16: lane group g supplies
// 8 of the 32 products' operands for A and B alike, so any assignment of
// k to those 8 slots that A and B share gives the same sum.  Here slots
// 0..3 are k-step kt's k = 4g..4g+3 and slots 4..7 k-step kt+1's: the B
// operand is two C-layout tiles side by side (no data movement), the A
// operand two 8-byte pieces of a weight row.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
__device__ inline bf16x8_t cat8(bf16x4_t lo, bf16x4_t hi) {
    typedef short s16x8_t __attribute__((ext_vector_type(8)));
    const s16x8_t v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_bit_cast(bf16x8_t, v);
}
__device__ inline f32x4_t mfma32(bf16x8_t a, bf16x8_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// A operand fragment of W (bf16 row-major, row stride `ld`): rows 16 mt ..,
// columns 16 kt + (l/16)*4 .. +3.
__device__ inline bf16x4_t frag_a(const uint16_t* W, int ld, int mt, int kt, int lane) {
    return *reinterpret_cast<const bf16x4_t*>(W + (16 * mt + (lane & 15)) * ld + 16 * kt + (lane >> 4) * 4);
}
// The A operand of k-steps kt and kt+1 together (mfma32).
__device__ inline bf16x8_t frag_a2(const uint16_t* W, int ld, int mt, int kt, int lane) {
    const uint16_t* p = W + (16 * mt + (lane & 15)) * ld + 16 * kt + (lane >> 4) * 4;
    return cat8(*reinterpret_cast<const bf16x4_t*>(p), *reinterpret_cast<const bf16x4_t*>(p + 16));
}

__device__ inline float logsigp(float u) {
    const float au = fabsf(u);
    return -au - 2.0f * __logf(1.0f + __expf(-au));
}

__device__ inline float tanh_f(float x) {
    const float e = __expf(-2.0f * fabsf(x));
    const float t = (1.0f - e) / (1.0f + e);
    return copysignf(t, x);
}

__device__ inline float group_sum(float v) {
    v += __shfl_xor(v, 16, kWave);
    v += __shfl_xor(v, 32, kWave);
    return v;
}

struct FlowArgs {
    int D, L;
    uint32_t amask[kFMaxL];       // conditioning set of each layer (bit d)
    const float* P;               // FP32 master parameters [L][kPLayer]
    const uint16_t* H;            // bf16 weights [L][kHLayer]
    const uint16_t* HT;           // bf16 transposes [L][kHLayer]
};

// One layer's weights and biases in LDS.
struct LayerLds {
    uint16_t W[kHLayer];
    float b2[kFH], b3[kFO];
};

__device__ inline void load_layer(const FlowArgs& f, int k, LayerLds* s, bool transposed, uint16_t* Wt) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(f.H + (size_t)k * kHLayer);
    uint32_t* dst = reinterpret_cast<uint32_t*>(s->W);
    for (int i = threadIdx.x; i < kHLayer / 2; i += blockDim.x) dst[i] = src[i];
    if (transposed) {
        const uint32_t* st = reinterpret_cast<const uint32_t*>(f.HT + (size_t)k * kHLayer);
        uint32_t* dt = reinterpret_cast<uint32_t*>(Wt);
        for (int i = threadIdx.x; i < kHLayer / 2; i += blockDim.x) dt[i] = st[i];
    }
    const float* p = f.P + (size_t)k * kPLayer;
    for (int i = threadIdx.x; i < kFH; i += blockDim.x) s->b2[i] = p[kPB2 + i];
    for (int i = threadIdx.x; i < kFO; i += blockDim.x) s->b3[i] = p[kPB3 + i];
}

// The conditioning input a of one tile as the B operand (dims (l/16)*4 + j).
__device__ inline bf16x4_t cond_input(const f32x4_t u, uint32_t amask, int D, int lane) {
    f32x4_t a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int d = (lane >> 4) * 4 + j;
        a[j] = (d < D && ((amask >> d) & 1u)) ? u[j] : (d == kFI - 1 ? 1.0f : 0.0f);
    }
    return pack_bf(a);
}

// Layer 1 of the conditioner MLP: h1 = relu(W1 a) (bf16 fragments, output
// tile mt = k-step mt of layer 2) and its ReLU mask (bit 4 mt + i:
// pre-activation > 0).
__device__ inline uint32_t layer1(const LayerLds* s, bf16x4_t a, int lane, bf16x4_t* h1) {
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    uint32_t m1 = 0;
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4_t c = mfma16(frag_a(s->W + kHW1, kS1, mt, 0, lane), a, zero);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            m1 |= (c[i] > 0.f ? 1u : 0u) << (4 * mt + i);
            c[i] = fmaxf(c[i], 0.f);
        }
        h1[mt] = pack_bf(c);
    }
    return m1;
}

// The conditioner MLP of one tile: o[0] = raw s, o[1] = t rows (l/16)*4+i.
// Optionally keeps h2 (bf16 fragments) and the ReLU masks of h1 and h2.
struct MlpKeep {
    bf16x4_t h2[8];
    uint32_t m1, m2;   // bit 4 mt + i: pre-activation > 0
};

template <bool KEEP>
__device__ inline void mlp(const LayerLds* s, bf16x4_t a, int lane, f32x4_t* o, MlpKeep* keep) {
    bf16x4_t h1[8];
    uint32_t m2 = 0;
    const uint32_t m1 = layer1(s, a, lane, h1);
    // h2: the eight output tiles accumulate side by side (independent MFMA
    // chains), k-step by k-step
    f32x4_t acc[8];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[mt][i] = s->b2[16 * mt + (lane >> 4) * 4 + i];
#pragma unroll
    for (int kt = 0; kt < 8; kt += 2) {
        const bf16x8_t b = cat8(h1[kt], h1[kt + 1]);
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = mfma32(frag_a2(s->W + kHW2, kS2, mt, kt, lane), b, acc[mt]);
        __builtin_amdgcn_sched_barrier(0);
    }
    bf16x4_t h2[8];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            m2 |= (acc[mt][i] > 0.f ? 1u : 0u) << (4 * mt + i);
            acc[mt][i] = fmaxf(acc[mt][i], 0.f);
        }
        h2[mt] = pack_bf(acc[mt]);
    }
    f32x4_t o0, o1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        o0[i] = s->b3[(lane >> 4) * 4 + i];
        o1[i] = s->b3[16 + (lane >> 4) * 4 + i];
    }
#pragma unroll
    for (int kt = 0; kt < 8; kt += 2) {
        const bf16x8_t b = cat8(h2[kt], h2[kt + 1]);
        o0 = mfma32(frag_a2(s->W + kHW3, kS3, 0, kt, lane), b, o0);
        o1 = mfma32(frag_a2(s->W + kHW3, kS3, 1, kt, lane), b, o1);
    }
    o[0] = o0;
    o[1] = o1;
    if (KEEP) {
#pragma unroll
        for (int q = 0; q < 8; ++q) keep->h2[q] = h2[q];
        keep->m1 = m1;
        keep->m2 = m2;
    }
}

// ------------------------------------------------------------- sampling ----
constexpr int kFlowThreads = 1024;
constexpr int kFlowTiles = 4;        // sample tiles (16 samples) per wave per pass
constexpr int kFlowWaves = kFlowThreads / 64;
constexpr int kFlowPass = kFlowWaves * kFlowTiles * 16;  // samples per workgroup pass

struct FlowFcn {
    int kind;            // VGX_FCN_GAUSSMIX or VGX_FCN_DIAGRIDGE
    int ncomp;
    const float* A;
    const float* B;
    const float* c;      // GAUSSMIX: [ncomp][D]; DIAGRIDGE: [ncomp]
};

struct SampleArgs {
    uint64_t n;          // samples
    uint64_t base;       // first sample index (Philox counter)
    uint32_t seed_lo, seed_hi, iteration;
    double* part;        // [gridDim.x][4]: DD sum f J, DD sum (f J)^2
    float* xout;         // optional [n][16] x, and fout [n] = f J
    float* fout;
    int32_t* status;
};

// f at x for the sample of this lane (dims (l/16)*4 + i), all four lanes of
// the sample cooperating; every lane of the sample returns f.
__device__ inline float flow_fcn(const FlowFcn& F, const float* Lc, f32x4_t x, int D, int lane) {
    const int g = lane >> 4;
    float acc = 0.f;
    if (F.kind == VGX_FCN_DIAGRIDGE) {
        float s1 = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * g + i < D) s1 += x[i];
        const float xbar = group_sum(s1) / (float)D;
        float p = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * g + i < D) p += (x[i] - xbar) * (x[i] - xbar);
        const float P = group_sum(p);
        const float fd = (float)D;
        for (int k = g; k < F.ncomp; k += 4) {
            const float t = xbar - Lc[2 * F.ncomp + k];
            acc = __builtin_fmaf(Lc[k], __expf(-Lc[F.ncomp + k] * __builtin_fmaf(fd * t, t, P)), acc);
        }
        return group_sum(acc);
    }
    for (int k = 0; k < F.ncomp; ++k) {
        float r = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * g + i < D) {
                const float t = x[i] - Lc[2 * F.ncomp + k * D + 4 * g + i];
                r = __builtin_fmaf(t, t, r);
            }
        r = group_sum(r);
        acc = __builtin_fmaf(Lc[k], __expf(-Lc[F.ncomp + k] * r), acc);
    }
    return acc;
}

// Per-wave tile state in LDS (kFlowTiles tiles: u as float4 per lane, and
// the log-Jacobian), so that tiles are indexed at run time without scratch.
struct TileState {
    f32x4_t u[kFlowTiles][64];
    float l[kFlowTiles][64];
};
constexpr size_t kStateLds = kFlowWaves * sizeof(TileState);

__global__ __launch_bounds__(kFlowThreads) void flow_sample_kernel(FlowArgs f, FlowFcn F, SampleArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    LayerLds* L = reinterpret_cast<LayerLds*>(smem);
    TileState* TS = reinterpret_cast<TileState*>(smem + ((sizeof(LayerLds) + 15) & ~(size_t)15)) +
                    (threadIdx.x >> 6);
    float* Lc = reinterpret_cast<float*>(smem + ((sizeof(LayerLds) + 15) & ~(size_t)15) + kStateLds);
    __shared__ double red[kFlowWaves * 4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4;
    const int ncp = F.kind == VGX_FCN_DIAGRIDGE ? 3 * F.ncomp : F.ncomp * (2 + f.D);
    for (int i = threadIdx.x; i < ncp; i += blockDim.x)
        Lc[i] = i < F.ncomp ? F.A[i] : i < 2 * F.ncomp ? F.B[i - F.ncomp] : F.c[i - 2 * F.ncomp];
    DD s1{0, 0}, s2{0, 0};
    for (uint64_t pb = (uint64_t)blockIdx.x * kFlowPass; pb < a.n; pb += (uint64_t)gridDim.x * kFlowPass) {
#pragma unroll 1
        for (int q = 0; q < kFlowTiles; ++q) {
            const uint64_t sidx = a.base + pb + (uint64_t)(wave * kFlowTiles + q) * 16 + (lane & 15);
            const U4 w = philox4x32_10_dev(U4{(uint32_t)sidx, (uint32_t)(sidx >> 32), a.iteration, (uint32_t)g},
                                       a.seed_lo, a.seed_hi);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
            float l = 0.f;
            f32x4_t uq;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float z = (float)(ws[i] >> 8) * (1.0f / 16777216.0f) + (1.0f / 33554432.0f);
                const float uu = __logf(z) - __logf(1.0f - z);
                uq[i] = uu;
                if (4 * g + i < f.D) l -= logsigp(uu);
            }
            TS->u[q][lane] = uq;
            TS->l[q][lane] = l;
        }
        for (int k = 0; k < f.L; ++k) {
            __syncthreads();
            load_layer(f, k, L, false, nullptr);
            __syncthreads();
            const uint32_t am = f.amask[k];
#pragma unroll 1
            for (int q = 0; q < kFlowTiles; ++q) {
                f32x4_t uq = TS->u[q][lane];
                float l = TS->l[q][lane];
                f32x4_t o[2];
                mlp<false>(L, cond_input(uq, am, f.D, lane), lane, o, nullptr);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int d = 4 * g + i;
                    if (d < f.D && !((am >> d) & 1u)) {
                        const float sv = kSMax * tanh_f(o[0][i] / kSMax);
                        uq[i] = uq[i] * __expf(sv) + o[1][i];
                        l += sv;
                    }
                }
                TS->u[q][lane] = uq;
                TS->l[q][lane] = l;
            }
        }
#pragma unroll 1
        for (int q = 0; q < kFlowTiles; ++q) {
            const f32x4_t uq = TS->u[q][lane];
            f32x4_t x;
            float l = TS->l[q][lane];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                x[i] = 1.0f / (1.0f + __expf(-uq[i]));
                if (4 * g + i < f.D) l += logsigp(uq[i]);
            }
            const float LJ = group_sum(l);
            const float fv = flow_fcn(F, Lc, x, f.D, lane);
            const uint64_t sl = pb + (uint64_t)(wave * kFlowTiles + q) * 16 + (lane & 15);
            const bool live = sl < a.n;
            // f = 0 where the map stretches enough for J to overflow is 0
            const double fj = (live && fv != 0.f) ? (double)fv * exp((double)LJ) : 0.0;
            if (fv != fv && live && a.status) atomicOr(a.status, (int32_t)VGX_ENAN);
            if (g == 0 && live) {
                dd_add(s1, fj);
                dd_add(s2, fj * fj);
            }
            if (a.xout && live) {
                *reinterpret_cast<f32x4_t*>(a.xout + sl * 16 + 4 * g) = x;
                if (g == 0) a.fout[sl] = (float)fj;
            }
        }
    }
    s1 = wave_sum_dd(s1);
    s2 = wave_sum_dd(s2);
    if (lane == 0) {
        red[4 * wave] = s1.hi;
        red[4 * wave + 1] = s1.lo;
        red[4 * wave + 2] = s2.hi;
        red[4 * wave + 3] = s2.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t1{0, 0}, t2{0, 0};
        for (int w = 0; w < kFlowWaves; ++w) {
            t1 = dd_sum(t1, DD{red[4 * w], red[4 * w + 1]});
            t2 = dd_sum(t2, DD{red[4 * w + 2], red[4 * w + 3]});
        }
        double* o = a.part + 4 * (size_t)blockIdx.x;
        o[0] = t1.hi; o[1] = t1.lo; o[2] = t2.hi; o[3] = t2.lo;
    }
}

// The integrand at given points x [n][16] (training data): f [n].
__global__ __launch_bounds__(256) void flow_eval_kernel(FlowFcn F, int D, const float* x, uint64_t n, float* fout) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* Lc = reinterpret_cast<float*>(smem);
    const int ncp = F.kind == VGX_FCN_DIAGRIDGE ? 3 * F.ncomp : F.ncomp * (2 + D);
    for (int i = threadIdx.x; i < ncp; i += blockDim.x)
        Lc[i] = i < F.ncomp ? F.A[i] : i < 2 * F.ncomp ? F.B[i - F.ncomp] : F.c[i - 2 * F.ncomp];
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint64_t ntiles = (n + 15) / 16;
    for (uint64_t t = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < ntiles;
         t += (uint64_t)gridDim.x * (blockDim.x >> 6)) {
        const uint64_t sl = t * 16 + (lane & 15);
        f32x4_t xv = {0.5f, 0.5f, 0.5f, 0.5f};
        if (sl < n) xv = *reinterpret_cast<const f32x4_t*>(x + sl * 16 + 4 * g);
        const float fv = flow_fcn(F, Lc, xv, D, lane);
        if (g == 0 && sl < n) fout[sl] = fv;
    }
}

// -------------------------------------------------------------- inverse ----
struct InvArgs {
    uint64_t n;
    const float* x;      // [n][16]
    float* saved;        // [L][n][16]: input of each layer
    float* u0;           // [n][16]
    float* logq;         // [n]
};

__global__ __launch_bounds__(kFlowThreads) void flow_inverse_kernel(FlowArgs f, InvArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    LayerLds* L = reinterpret_cast<LayerLds*>(smem);
    TileState* TS = reinterpret_cast<TileState*>(smem + ((sizeof(LayerLds) + 15) & ~(size_t)15)) +
                    (threadIdx.x >> 6);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4;
    for (uint64_t pb = (uint64_t)blockIdx.x * kFlowPass; pb < a.n; pb += (uint64_t)gridDim.x * kFlowPass) {
#pragma unroll 1
        for (int q = 0; q < kFlowTiles; ++q) {
            const uint64_t sl = pb + (uint64_t)(wave * kFlowTiles + q) * 16 + (lane & 15);
            f32x4_t xv = {0.5f, 0.5f, 0.5f, 0.5f};
            if (sl < a.n) xv = *reinterpret_cast<const f32x4_t*>(a.x + sl * 16 + 4 * g);
            float l = 0.f;
            f32x4_t uq;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float xc = fminf(fmaxf(xv[i], 1e-7f), 1.0f - 1e-7f);
                const float uu = __logf(xc) - __logf(1.0f - xc);
                uq[i] = (4 * g + i < f.D) ? uu : 0.f;
                if (4 * g + i < f.D) l -= logsigp(uu);
            }
            TS->u[q][lane] = uq;
            TS->l[q][lane] = l;
        }
        for (int k = f.L - 1; k >= 0; --k) {
            __syncthreads();
            load_layer(f, k, L, false, nullptr);
            __syncthreads();
            const uint32_t am = f.amask[k];
#pragma unroll 1
            for (int q = 0; q < kFlowTiles; ++q) {
                const uint64_t sl = pb + (uint64_t)(wave * kFlowTiles + q) * 16 + (lane & 15);
                f32x4_t uq = TS->u[q][lane];
                float l = TS->l[q][lane];
                if (sl < a.n) *reinterpret_cast<f32x4_t*>(a.saved + ((size_t)k * a.n + sl) * 16 + 4 * g) = uq;
                f32x4_t o[2];
                mlp<false>(L, cond_input(uq, am, f.D, lane), lane, o, nullptr);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int d = 4 * g + i;
                    if (d < f.D && !((am >> d) & 1u)) {
                        const float sv = kSMax * tanh_f(o[0][i] / kSMax);
                        uq[i] = (uq[i] - o[1][i]) * __expf(-sv);
                        l -= sv;
                    }
                }
                TS->u[q][lane] = uq;
                TS->l[q][lane] = l;
            }
        }
#pragma unroll 1
        for (int q = 0; q < kFlowTiles; ++q) {
            const uint64_t sl = pb + (uint64_t)(wave * kFlowTiles + q) * 16 + (lane & 15);
            const f32x4_t uq = TS->u[q][lane];
            float l = TS->l[q][lane];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * g + i < f.D) l += logsigp(uq[i]);
            const float LQ = group_sum(l);
            if (sl < a.n) {
                *reinterpret_cast<f32x4_t*>(a.u0 + sl * 16 + 4 * g) = uq;
                if (g == 0) a.logq[sl] = LQ;
            }
        }
    }
}

// ---------------------------------------------------------- loss weights ----
// log c_i = log w_i + 2 log|f_i| - log q_i (w given by its logarithm, so
// weights 1/(n q) of the flow's own samples cannot overflow); c normalised
// to sum 1 (two launches: the maximum, then the sum of exp(log c - max)).
// Terms that are not finite count as 0.
// pf, pq: exponents of |f| and q -- (2, 1) the variance loss, (1, 0) the
// forward KL divergence (maximum likelihood of the density |f|/I).
__device__ inline float lossw_log(const float* lw, const float* fv, const float* logq, uint64_t i, float pf,
                                  float pq) {
    const float fa = fabsf(fv[i]);
    if (!(fa > 0.f)) return -INFINITY;
    const float v = lw[i] + pf * __logf(fa) - pq * logq[i];
    return (v == v && fabsf(v) < INFINITY) ? v : -INFINITY;
}

__device__ inline uint32_t ord_f(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float unord_f(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// One atomic per workgroup (the maximum is exact in any order).
__global__ __launch_bounds__(256) void flow_lossw_max_kernel(const float* w, const float* fv, const float* logq,
                                                             uint64_t n, float pf, float pq, uint32_t* mx) {
    __shared__ float red[4];
    float m = -INFINITY;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = fmaxf(m, lossw_log(w, fv, logq, i, pf, pq));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, kWave));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        if (m > -INFINITY) atomicMax(mx, ord_f(m));
    }
}

__global__ void flow_lossw_sum_kernel(const float* w, const float* fv, const float* logq, uint64_t n,
                                      float pf, float pq, const uint32_t* mx, float* c, double* part) {
    __shared__ double red[16];
    const bool none = *mx == 0u;  // no sample with f != 0
    const float m = none ? 0.f : unord_f(*mx);
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float v = none ? 0.f : __expf(lossw_log(w, fv, logq, i, pf, pq) - m);
        c[i] = v;
        s += v;
    }
    s = wave_sum_fixed(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
        part[blockIdx.x] = t;
    }
}

// The sum of the loss weights' partial sums, in a fixed order (thread t
// adds parts t, t + 256, ... in order; then the threads' sums in order).
__global__ __launch_bounds__(256) void flow_csum_kernel(const double* part, int nparts, double* out) {
    __shared__ double red[256];
    double t = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) t += part[i];
    red[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double u = 0.0;
        for (int i = 0; i < 256; ++i) u += red[i];
        *out = u;
    }
}

// ------------------------------------------------------------- backward ----
struct BwdArgs {
    uint64_t n;
    int k;                    // layer
    const float* saved;       // [L][n][16]
    const float* u0;          // [n][16]
    const float* c;           // [n] unnormalised weights
    const double* csum;       // the sum of c (flow_csum_kernel)
    float* G;                 // [n][16] dL/du (in: layer output; out: layer input)
    float* part;              // [gridDim.x][kPLayer] each workgroup's weight-gradient sums
};

// Backward of one layer together with its weight gradients.  Per round each
// wave takes one 16-sample tile (tile (round * gridDim.x + block) * 8 +
// wave), recomputes the layer's MLP, backpropagates through the coupling and
// the MLP, and the eight waves stage their GEMM operands in LDS (element
// (row, sample) of a wave's slot at row * 16 + sample) in three phases:
//   1: dout (32 rows) | h2 (128)   -> dW3, b3
//   2: dh2 (128) | h1 (128)        -> dW2, b2 (h1 recomputed from a: 8 small MFMAs
//                                     instead of 16 registers held through the
//                                     backward GEMMs)
//   3: dh1 (128) | a (16)          -> dW1
// After each staging every wave adds its share of dW over the eight staged
// tiles, two tiles per v_mfma_f32_16x16x32_bf16 (k = the samples of both);
// an all-ones operand gives the bias sums.  The sums stay in registers over
// the rounds and are written once per workgroup: the operands never leave
// the CU.  The tile -> workgroup assignment is static, so the sums are
// reproducible.  Tiles past the end run on zero inputs: their dout is 0, so
// they add nothing.
constexpr int kBwdThreads = 512;
constexpr int kBwdWaves = kBwdThreads / 64;
constexpr int kSlotRows = 2 * kFH;         // the largest phase: dh2 | h1
constexpr int kSlot = kSlotRows * 16;      // bf16 elements per wave slot
constexpr int kNo3 = 2 * 9;                // dW3 output tiles: 2 row tiles x (8 column tiles + bias)

// The C-layout tile v (rows row0 + (l/16)*4 + i, sample l%16) into a slot.
// Lane (g, 4m + t) holds rows 4g..4g+3 of sample 4m + t; a 4x4 transpose of
// the 16-bit values within the lane quad (DPP) leaves it row 4g + t of
// samples 4m..4m+3, which it stores as one 8-byte write (at the permuted
// chunk position frag_op reads).
__device__ inline void stage(uint16_t* slot, int row0, const bf16x4_t v, int lane) {
    typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
    const u32x2_t d = __builtin_bit_cast(u32x2_t, v);
    const int t = lane & 3;
    // dwords with lane t^2: lanes 0,1 keep (x0,x1) and take (x0,x1) of lane
    // t+2; lanes 2,3 take (x2,x3) of lane t-2 and keep (x2,x3)
    const uint32_t p0 = (uint32_t)__builtin_amdgcn_mov_dpp((int)d[0], 0x4E, 0xF, 0xF, false);
    const uint32_t p1 = (uint32_t)__builtin_amdgcn_mov_dpp((int)d[1], 0x4E, 0xF, 0xF, false);
    const uint32_t a0 = t < 2 ? d[0] : p1, a1 = t < 2 ? p0 : d[1];
    // 16-bit halves with lane t^1: even lanes lo(own) | lo(odd) << 16, odd
    // lanes hi(even) | hi(own) << 16
    const uint32_t q0 = (uint32_t)__builtin_amdgcn_mov_dpp((int)a0, 0xB1, 0xF, 0xF, false);
    const uint32_t q1 = (uint32_t)__builtin_amdgcn_mov_dpp((int)a1, 0xB1, 0xF, 0xF, false);
    const uint32_t sel = (t & 1) ? 0x03020706u : 0x05040100u;
    u32x2_t r;
    r[0] = __builtin_amdgcn_perm(q0, a0, sel);
    r[1] = __builtin_amdgcn_perm(q1, a1, sel);
    const int row = row0 + 4 * (lane >> 4) + t;
    *reinterpret_cast<u32x2_t*>(slot + row * 16 + 4 * (((lane >> 2) & 3) ^ ((row >> 2) & 3))) = r;
}
// Operand fragment of slot rows row0.. (A and B alike: row row0 + l%16,
// samples (l/16)*4 .. +3).  In a slot row the four 4-sample chunks are
// permuted, chunk c at position c ^ ((row >> 2) & 3): the 16 rows that one
// fragment read (and the 16 that one staging store) touches then fall in
// distinct LDS banks, where the plain 32-byte rows put every fourth row in
// the same banks.
__device__ inline bf16x4_t frag_op(const uint16_t* slot, int row0, int lane) {
    const int row = row0 + (lane & 15);
    return *reinterpret_cast<const bf16x4_t*>(slot + row * 16 + 4 * ((lane >> 4) ^ ((row >> 2) & 3)));
}
// Rows row0.. of staged tiles j and j+1 as one k = 32 operand (A and B
// alike: slots 0..3 tile j's samples (l/16)*4.., slots 4..7 tile j+1's).
__device__ inline bf16x8_t frag_pair(const uint16_t* S, int j, int row0, int lane) {
    const uint16_t* s0 = S + j * kSlot;
    return cat8(frag_op(s0, row0, lane), frag_op(s0 + kSlot, row0, lane));
}

constexpr size_t kLdsLayer = (sizeof(LayerLds) + 15) & ~(size_t)15;
constexpr size_t kLdsBwd = kLdsLayer + sizeof(uint16_t) * kHLayer + sizeof(uint16_t) * (size_t)kBwdWaves * kSlot;
static_assert(kLdsBwd <= 160 * 1024, "backward LDS plan fits a CU");

__global__ __launch_bounds__(kBwdThreads) void flow_backward_kernel(FlowArgs f, BwdArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    LayerLds* L = reinterpret_cast<LayerLds*>(smem);
    uint16_t* Wt = reinterpret_cast<uint16_t*>(smem + kLdsLayer);
    uint16_t* S = Wt + kHLayer;
    // the wave index as a wave-uniform (SGPR) value: the output tiles a wave
    // owns and its slot address are then scalar
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, g = lane >> 4;
    uint16_t* my = S + w * kSlot;
    const int k = a.k;
    load_layer(f, k, L, true, Wt);
    __syncthreads();
    const double csum = *a.csum;
    const float inv_c = csum > 0 ? (float)(1.0 / csum) : 0.f;
    const uint32_t am = f.amask[k];
    const uint64_t ntiles = (a.n + 15) / 16;
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    f32x4_t acc3[3], acc2[9], acc1 = zero;
#pragma unroll
    for (int q = 0; q < 9; ++q) acc2[q] = zero;
#pragma unroll
    for (int q = 0; q < 3; ++q) acc3[q] = zero;
    // all-ones operand (1 in column 0): row sums, the bias gradients
    bf16x4_t ones;
#pragma unroll
    for (int j = 0; j < 4; ++j) ones[j] = (lane & 15) == 0 ? (short)0x3F80 : (short)0;
    const bf16x8_t ones8 = cat8(ones, ones);
    // The inputs of a wave's next tile (saved u, c, and G or u0) are loaded
    // while the current tile computes.
    auto fetch = [&](uint64_t tl, f32x4_t& u, f32x4_t& gr, float& c) {
        const uint64_t s2 = tl * 16 + (lane & 15);
        u = zero;
        gr = zero;
        c = 0.f;
        if (tl < ntiles && s2 < a.n) {
            u = *reinterpret_cast<const f32x4_t*>(a.saved + ((size_t)k * a.n + s2) * 16 + 4 * g);
            c = a.c[s2];
            gr = *reinterpret_cast<const f32x4_t*>((k == 0 ? a.u0 : a.G) + s2 * 16 + 4 * g);
        }
    };
    const uint64_t per_round = (uint64_t)gridDim.x * kBwdWaves;
    const uint64_t rounds = (ntiles + per_round - 1) / per_round;
    uint64_t tile = (uint64_t)blockIdx.x * kBwdWaves + w;
    f32x4_t nx_u, nx_g;
    float nx_c;
    fetch(tile, nx_u, nx_g, nx_c);
    for (uint64_t rd = 0; rd < rounds; ++rd, tile += per_round) {
        const uint64_t sl = tile * 16 + (lane & 15);
        const bool live = sl < a.n;
        const f32x4_t uin = nx_u, graw = nx_g;
        const float cw = nx_c * inv_c;
        fetch(tile + per_round, nx_u, nx_g, nx_c);
        f32x4_t G = graw;
        if (k == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                G[i] = (live && 4 * g + i < f.D) ? -cw * (1.0f - 2.0f / (1.0f + __expf(-graw[i]))) : 0.f;
        }
        const bf16x4_t ain = cond_input(uin, am, f.D, lane);
        f32x4_t o[2];
        MlpKeep kp;
        mlp<true>(L, ain, lane, o, &kp);
        // coupling backward (inverse direction)
        f32x4_t dout0, dout1, Gin;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int d = 4 * g + i;
            dout0[i] = 0.f;
            dout1[i] = 0.f;
            Gin[i] = G[i];
            if (d < f.D && !((am >> d) & 1u)) {
                const float th = tanh_f(o[0][i] / kSMax);
                const float sv = kSMax * th;
                const float es = __expf(-sv);
                const float uout = (uin[i] - o[1][i]) * es;
                const float ds = -G[i] * uout + cw;
                dout0[i] = live ? ds * (1.0f - th * th) : 0.f;
                dout1[i] = live ? -G[i] * es : 0.f;
                Gin[i] = G[i] * es;
            }
        }
        const bf16x4_t do0 = pack_bf(dout0), do1 = pack_bf(dout1);
        // The staging and operand addresses of the phases below are formed
        // from this copy of the lane index when used (immediate offsets of the
        // LDS instructions): held from round to round they would take about
        // forty registers.
        int ln = lane;
        asm volatile("" : "+v"(ln));
        // ---- phase 1: dout | h2 -> dW3, b3 (output tile o = wave + 8 j)
        stage(my, 0, do0, ln);
        stage(my, 16, do1, ln);
#pragma unroll
        for (int q = 0; q < 8; ++q) stage(my, 32 + 16 * q, kp.h2[q], ln);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kBwdWaves; j += 2) {
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int ot = w + kBwdWaves * e;
                if (ot < kNo3) {
                    const int q = ot % 9;
                    acc3[e] = mfma32(frag_pair(S, j, 16 * (ot / 9), ln),
                                     q < 8 ? frag_pair(S, j, 32 + 16 * q, ln) : ones8, acc3[e]);
                }
            }
        }
        // dh2 = W3^T dout (relu mask h2); dh1 = W2^T dh2 (mask h1); da = W1^T dh1
        bf16x4_t dh2[8];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
            f32x4_t cc = mfma32(frag_a2(Wt + kTW3, kT3, mt, 0, lane), cat8(do0, do1), zero);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (!((kp.m2 >> (4 * mt + i)) & 1u)) cc[i] = 0.f;
            dh2[mt] = pack_bf(cc);
        }
        f32x4_t acc[8];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = zero;
#pragma unroll
        for (int kt = 0; kt < 8; kt += 2) {
            const bf16x8_t b = cat8(dh2[kt], dh2[kt + 1]);
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) acc[mt] = mfma32(frag_a2(Wt + kTW2, kT2, mt, kt, lane), b, acc[mt]);
            __builtin_amdgcn_sched_barrier(0);
        }
        bf16x4_t dh1[8];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (!((kp.m1 >> (4 * mt + i)) & 1u)) acc[mt][i] = 0.f;
            dh1[mt] = pack_bf(acc[mt]);
        }
        f32x4_t da = zero;
#pragma unroll
        for (int kt = 0; kt < 8; kt += 2) da = mfma32(frag_a2(Wt + kTW1, kT1, 0, kt, lane), cat8(dh1[kt], dh1[kt + 1]), da);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int d = 4 * g + i;
            if (d < f.D && ((am >> d) & 1u)) Gin[i] += da[i];
        }
        if (live) *reinterpret_cast<f32x4_t*>(a.G + sl * 16 + 4 * g) = Gin;
        __syncthreads();
        // ---- phase 2: dh2 | h1 -> dW2, b2
        {
            bf16x4_t h1[8];
            layer1(L, ain, lane, h1);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                stage(my, 16 * q, dh2[q], ln);
                stage(my, kFH + 16 * q, h1[q], ln);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kBwdWaves; j += 2) {
            const bf16x8_t av = frag_pair(S, j, 16 * w, ln);
#pragma unroll
            for (int q = 0; q < 8; ++q) acc2[q] = mfma32(av, frag_pair(S, j, kFH + 16 * q, ln), acc2[q]);
            acc2[8] = mfma32(av, ones8, acc2[8]);
        }
        __syncthreads();
        // ---- phase 3: dh1 | a -> dW1
#pragma unroll
        for (int q = 0; q < 8; ++q) stage(my, 16 * q, dh1[q], ln);
        stage(my, kFH, ain, ln);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kBwdWaves; j += 2) acc1 = mfma32(frag_pair(S, j, 16 * w, ln), frag_pair(S, j, kFH, ln), acc1);
        __syncthreads();
    }
    // C layout: lane l, i -> dW[16 r + (l/16)*4 + i][16 q + l%16]
    float* o = a.part + (size_t)blockIdx.x * kPLayer;
    const int cc = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int rr = 4 * g + i, r2 = 16 * w + rr;
#pragma unroll
        for (int q = 0; q < 8; ++q) o[kPW2 + r2 * kFH + 16 * q + cc] = acc2[q][i];
        if (cc == 0) o[kPB2 + r2] = acc2[8][i];
        o[kPW1 + r2 * kFI + cc] = acc1[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int ot = w + kBwdWaves * j;
            if (ot < kNo3) {
                const int q = ot % 9, row = 16 * (ot / 9) + rr;
                if (q < 8)
                    o[kPW3 + row * kFH + 16 * q + cc] = acc3[j][i];
                else if (cc == 0)
                    o[kPB3 + row] = acc3[j][i];
            }
        }
    }
}

// Sum of the workgroup partials into grad [kPLayer], in a fixed order: a
// workgroup takes 32 consecutive entries; its 8 part groups (thread j sums
// parts j, j + 8, ... in order) are added in group order.  Each wave reads
// 2 x 128 contiguous bytes per load, and 709 workgroups keep the loads of
// the whole device in flight.
constexpr int kRedCols = 32, kRedGroups = 8;

__global__ __launch_bounds__(kRedCols * kRedGroups) void flow_wgrad_reduce_kernel(const float* part, int nparts,
                                                                                  float* grad) {
    __shared__ double red[kRedGroups][kRedCols];
    const int il = threadIdx.x % kRedCols, j = threadIdx.x / kRedCols;
    const int i = blockIdx.x * kRedCols + il;
    double s = 0.0;
    if (i < kPLayer) {
        int p = j;
        for (; p + 3 * kRedGroups < nparts; p += 4 * kRedGroups) {
            const float v0 = part[(size_t)p * kPLayer + i];
            const float v1 = part[(size_t)(p + kRedGroups) * kPLayer + i];
            const float v2 = part[(size_t)(p + 2 * kRedGroups) * kPLayer + i];
            const float v3 = part[(size_t)(p + 3 * kRedGroups) * kPLayer + i];
            s += v0;
            s += v1;
            s += v2;
            s += v3;
        }
        for (; p < nparts; p += kRedGroups) s += part[(size_t)p * kPLayer + i];
    }
    red[j][il] = s;
    __syncthreads();
    if (j == 0 && i < kPLayer) {
        double t = 0.0;
        for (int q = 0; q < kRedGroups; ++q) t += red[q][il];
        grad[i] = (float)t;
    }
}

// ------------------------------------------------------------------ adam ----
struct AdamArgs {
    float* P;
    const float* grad;
    float* m;
    float* v;
    uint16_t* H;
    uint16_t* HT;
    int L;
    float lr, b1, b2, eps, c1, c2;   // c1 = 1 - b1^t, c2 = 1 - b2^t
    float clip;                       // gradient clip (max abs), 0 = none
};

__global__ void flow_adam_kernel(AdamArgs a) {
    const int total = a.L * kPLayer;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        float gr = a.grad[i];
        if (!(fabsf(gr) < INFINITY)) gr = 0.f;  // a non-finite component is skipped
        if (a.clip > 0) gr = fminf(fmaxf(gr, -a.clip), a.clip);
        const float m = a.b1 * a.m[i] + (1.f - a.b1) * gr;
        const float v = a.b2 * a.v[i] + (1.f - a.b2) * gr * gr;
        a.m[i] = m;
        a.v[i] = v;
        a.P[i] = a.P[i] - a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
    }
}

// bf16 copies (and transposes) of the weights after an update, in the
// padded layouts.
__global__ void flow_cast_kernel(const float* P, uint16_t* H, uint16_t* HT, int L) {
    constexpr int nlog = kFH * kFI + kFH * kFH + kFO * kFH;
    const int total = L * nlog;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int k = i / nlog, r = i % nlog;
        const float* p = P + (size_t)k * kPLayer;
        float v;
        int rh, rt;
        if (r < kFH * kFI) {                 // W1 [128][16] -> W1t [16][128]
            const int o = r / kFI, c = r % kFI;
            v = p[kPW1 + r];
            rh = kHW1 + o * kS1 + c;
            rt = kTW1 + c * kT1 + o;
        } else if (r < kFH * kFI + kFH * kFH) {  // W2 [128][128]
            const int q = r - kFH * kFI, o = q / kFH, c = q % kFH;
            v = p[kPW2 + q];
            rh = kHW2 + o * kS2 + c;
            rt = kTW2 + c * kT2 + o;
        } else {                             // W3 [32][128] -> W3t [128][32]
            const int q = r - kFH * kFI - kFH * kFH, o = q / kFH, c = q % kFH;
            v = p[kPW3 + q];
            rh = kHW3 + o * kS3 + c;
            rt = kTW3 + c * kT3 + o;
        }
        const uint16_t h = f2bf(v);
        H[(size_t)k * kHLayer + rh] = h;
        HT[(size_t)k * kHLayer + rt] = h;
    }
}

__global__ void flow_zero_u32_kernel(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// MFMA lane-layout probe: D = A B for 16x16 bf16 A, B given row-major.
__global__ void flow_mfma_probe_kernel(const uint16_t* A, const uint16_t* B, float* D) {
    const int lane = threadIdx.x & 63;
    bf16x4_t a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        a[j] = (short)A[(lane & 15) * 16 + (lane >> 4) * 4 + j];
        b[j] = (short)B[((lane >> 4) * 4 + j) * 16 + (lane & 15)];
    }
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    const f32x4_t c = mfma16(a, b, zero);
#pragma unroll
    for (int i = 0; i < 4; ++i) D[((lane >> 4) * 4 + i) * 16 + (lane & 15)] = c[i];
}

}  // namespace vgx

// ------------------------------------------------------------ host side ----
using namespace vgx;

struct vgx_flow {
    int D = 0, L = 0;
    uint32_t amask[kFMaxL] = {};
    float* P = nullptr;      // FP32 master [L][kPLayer]
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    uint16_t* H = nullptr;
    uint16_t* HT = nullptr;
    // per-batch scratch (grown on demand)
    Scratch saved, u0, logq, c, G, wpart, cpart, spart, misc, fpar;
    int step = 0;
};

static FlowArgs flow_args(const vgx_flow* fl) {
    FlowArgs f{};
    f.D = fl->D;
    f.L = fl->L;
    for (int k = 0; k < fl->L; ++k) f.amask[k] = fl->amask[k];
    f.P = fl->P;
    f.H = fl->H;
    f.HT = fl->HT;
    return f;
}

static size_t lds_layer() { return kLdsLayer; }
// backward: the layer, the transposed weights, the operand slots of the waves
static size_t lds_bwd() { return kLdsBwd; }

static int flow_cast(vgx_flow* fl, hipStream_t s) {
    hipLaunchKernelGGL(flow_cast_kernel, dim3(256), dim3(256), 0, s, (const float*)fl->P, fl->H, fl->HT, fl->L);
    return check_launch();
}

extern "C" int vgx_flow_destroy(vgx_flow* fl);

extern "C" int vgx_flow_create(int dim, int nlayers, const uint32_t* amask, vgx_flow** out) {
    if (!out || !amask || dim < 1 || dim > kFI - 1 || nlayers < 1 || nlayers > kFMaxL) return VGX_EINVAL;
    vgx_flow* fl = new vgx_flow();
    fl->D = dim;
    fl->L = nlayers;
    for (int k = 0; k < nlayers; ++k) fl->amask[k] = amask[k] & ((1u << dim) - 1u);
    const size_t np = (size_t)nlayers * kPLayer, nh = (size_t)nlayers * kHLayer;
    if (hipMalloc(&fl->P, np * 4) != hipSuccess || hipMalloc(&fl->m, np * 4) != hipSuccess ||
        hipMalloc(&fl->v, np * 4) != hipSuccess || hipMalloc(&fl->grad, np * 4) != hipSuccess ||
        hipMalloc(&fl->H, nh * 2) != hipSuccess || hipMalloc(&fl->HT, nh * 2) != hipSuccess) {
        delete fl;
        return VGX_ENOMEM;
    }
    if (hipMemset(fl->P, 0, np * 4) != hipSuccess || hipMemset(fl->m, 0, np * 4) != hipSuccess ||
        hipMemset(fl->v, 0, np * 4) != hipSuccess || hipMemset(fl->grad, 0, np * 4) != hipSuccess ||
        hipMemset(fl->H, 0, nh * 2) != hipSuccess || hipMemset(fl->HT, 0, nh * 2) != hipSuccess) {
        vgx_flow_destroy(fl);
        return VGX_EHIP;
    }
    if (hipFuncSetAttribute((const void*)flow_backward_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bwd()) != hipSuccess ||
        hipFuncSetAttribute((const void*)flow_inverse_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(lds_layer() + kStateLds)) != hipSuccess) {
        vgx_flow_destroy(fl);
        return VGX_EHIP;
    }
    *out = fl;
    return VGX_OK;
}

extern "C" int vgx_flow_destroy(vgx_flow* fl) {
    if (!fl) return VGX_OK;
    (void)hipDeviceSynchronize();
    void* ps[] = {fl->P, fl->m, fl->v, fl->grad, fl->H, fl->HT};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    Scratch* ss[] = {&fl->saved, &fl->u0, &fl->logq, &fl->c, &fl->G, &fl->wpart, &fl->cpart, &fl->spart,
                     &fl->misc, &fl->fpar};
    for (Scratch* q : ss) q->release();
    delete fl;
    return VGX_OK;
}

extern "C" int vgx_flow_nparams(const vgx_flow* fl, int64_t* n) {
    if (!fl || !n) return VGX_EINVAL;
    *n = (int64_t)fl->L * kPLayer;
    return VGX_OK;
}

extern "C" int vgx_flow_set_params(vgx_flow* fl, const float* host, void* stream) {
    if (!fl || !host) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    const size_t np = (size_t)fl->L * kPLayer;
    VGX_HIP(hipMemcpyAsync(fl->P, host, np * 4, hipMemcpyHostToDevice, s));
    VGX_HIP(hipMemsetAsync(fl->m, 0, np * 4, s));
    VGX_HIP(hipMemsetAsync(fl->v, 0, np * 4, s));
    fl->step = 0;
    int rc = flow_cast(fl, s);
    if (rc) return rc;
    VGX_HIP(hipStreamSynchronize(s));
    return VGX_OK;
}

extern "C" int vgx_flow_get_params(const vgx_flow* fl, float* host, float* grad_host, void* stream) {
    if (!fl) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    const size_t np = (size_t)fl->L * kPLayer;
    if (host) VGX_HIP(hipMemcpyAsync(host, fl->P, np * 4, hipMemcpyDeviceToHost, s));
    if (grad_host) VGX_HIP(hipMemcpyAsync(grad_host, fl->grad, np * 4, hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    return VGX_OK;
}

// Device integrand parameters converted to FP32 into fl->fpar.
static int flow_fcn_params(vgx_flow* fl, const vgx_integrand* fcn, FlowFcn* F, hipStream_t s) {
    if (!fcn || (fcn->kind != VGX_FCN_GAUSSMIX && fcn->kind != VGX_FCN_DIAGRIDGE) || fcn->ncomp < 1)
        return VGX_EINVAL;
    const int nc = fcn->ncomp;
    const int ncc = fcn->kind == VGX_FCN_GAUSSMIX ? nc * fl->D : nc;
    const size_t tot = (size_t)2 * nc + ncc;
    if (tot * 4 > 48 * 1024) return VGX_ERANGE;  // staged in LDS with the layer
    int rc = fl->fpar.ensure(tot * 4);
    if (rc) return rc;
    float* d = (float*)fl->fpar.ptr;
    std::vector<double> h(tot);
    VGX_HIP(hipMemcpyAsync(h.data(), fcn->A, nc * 8, hipMemcpyDeviceToHost, s));
    VGX_HIP(hipMemcpyAsync(h.data() + nc, fcn->B, nc * 8, hipMemcpyDeviceToHost, s));
    VGX_HIP(hipMemcpyAsync(h.data() + 2 * nc, fcn->c, ncc * 8, hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    std::vector<float> hf(tot);
    for (size_t i = 0; i < tot; ++i) hf[i] = (float)h[i];
    VGX_HIP(hipMemcpyAsync(d, hf.data(), tot * 4, hipMemcpyHostToDevice, s));
    VGX_HIP(hipStreamSynchronize(s));
    F->kind = fcn->kind;
    F->ncomp = nc;
    F->A = d;
    F->B = d + nc;
    F->c = d + 2 * nc;
    return VGX_OK;
}

extern "C" int vgx_flow_sample(vgx_flow* fl, const vgx_integrand* fcn, int64_t n, uint64_t base, uint64_t seed,
                               uint32_t iteration, double* sums, float* xout, float* fout, int32_t* status,
                               void* stream) {
    if (!fl || n < 1 || (!sums && (!xout || !fout))) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    FlowFcn F{};
    int rc = flow_fcn_params(fl, fcn, &F, s);
    if (rc) return rc;
    DeviceInfo info;
    if ((rc = device_info(&info))) return rc;
    const int blocks = (int)std::min<int64_t>((int64_t)info.cus, (n + kFlowPass - 1) / kFlowPass);
    if ((rc = fl->spart.ensure((size_t)blocks * 4 * 8))) return rc;
    SampleArgs a{};
    a.n = (uint64_t)n;
    a.base = base;
    a.seed_lo = (uint32_t)seed;
    a.seed_hi = (uint32_t)(seed >> 32);
    a.iteration = iteration;
    a.part = (double*)fl->spart.ptr;
    a.xout = xout;
    a.fout = fout;
    a.status = status;
    const size_t ncp = F.kind == VGX_FCN_DIAGRIDGE ? 3 * (size_t)F.ncomp : (size_t)F.ncomp * (2 + fl->D);
    const size_t lds = lds_layer() + kStateLds + ncp * 4;
    if (hipFuncSetAttribute((const void*)flow_sample_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        return VGX_EHIP;
    hipEvent_t e0, e1;
    if (timed_slot(g_timing_ctx, kTimedFlowSample, &e0, &e1))
        hipExtLaunchKernelGGL(flow_sample_kernel, dim3(blocks), dim3(kFlowThreads), (uint32_t)lds, s, e0, e1, 0,
                              flow_args(fl), F, a);
    else
        hipLaunchKernelGGL(flow_sample_kernel, dim3(blocks), dim3(kFlowThreads), lds, s, flow_args(fl), F, a);
    if ((rc = check_launch())) return rc;
    if (!sums) return VGX_OK;
    std::vector<double> h((size_t)blocks * 4);
    VGX_HIP(hipMemcpyAsync(h.data(), a.part, h.size() * 8, hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    DD t1{0, 0}, t2{0, 0};
    for (int b = 0; b < blocks; ++b) {
        t1 = dd_sum(t1, DD{h[4 * b], h[4 * b + 1]});
        t2 = dd_sum(t2, DD{h[4 * b + 2], h[4 * b + 3]});
    }
    sums[0] = t1.hi + t1.lo;
    sums[1] = t2.hi + t2.lo;
    return VGX_OK;
}

static int flow_inverse(vgx_flow* fl, const float* x, int64_t n, hipStream_t s) {
    int rc;
    const size_t nn = (size_t)n;
    if ((rc = fl->saved.ensure((size_t)fl->L * nn * 64))) return rc;
    if ((rc = fl->u0.ensure(nn * 64))) return rc;
    if ((rc = fl->logq.ensure(nn * 4))) return rc;
    DeviceInfo info;
    if ((rc = device_info(&info))) return rc;
    const int blocks = (int)std::min<int64_t>((int64_t)info.cus * 2, (n + kFlowPass - 1) / kFlowPass);
    InvArgs a{(uint64_t)n, x, (float*)fl->saved.ptr, (float*)fl->u0.ptr, (float*)fl->logq.ptr};
    hipLaunchKernelGGL(flow_inverse_kernel, dim3(blocks), dim3(kFlowThreads), lds_layer() + kStateLds, s,
                       flow_args(fl), a);
    return check_launch();
}

extern "C" int vgx_flow_logq(vgx_flow* fl, const float* x, int64_t n, float* logq, float* u0, void* stream) {
    if (!fl || !x || !logq || n < 1) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    int rc = flow_inverse(fl, x, n, s);
    if (rc) return rc;
    VGX_HIP(hipMemcpyAsync(logq, fl->logq.ptr, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    if (u0) VGX_HIP(hipMemcpyAsync(u0, fl->u0.ptr, (size_t)n * 64, hipMemcpyDeviceToDevice, s));
    return VGX_OK;
}

extern "C" int vgx_flow_grad(vgx_flow* fl, const float* x, const float* f, const float* logw, int64_t n,
                             int loss, void* stream) {
    if (!fl || !x || !f || !logw || n < 1) return VGX_EINVAL;
    if (loss != VGX_FLOW_LOSS_VARIANCE && loss != VGX_FLOW_LOSS_KL && loss != VGX_FLOW_LOSS_SQUARED) return VGX_EINVAL;
    const float* w = logw;
    hipStream_t s = (hipStream_t)stream;
    int rc = flow_inverse(fl, x, n, s);
    if (rc) return rc;
    const size_t nn = (size_t)n;
    const uint64_t ntiles = (nn + 15) / 16;
    const int nc = 1024;
    if ((rc = fl->c.ensure(nn * 4))) return rc;
    if ((rc = fl->G.ensure(nn * 64))) return rc;
    if ((rc = fl->cpart.ensure(nc * 8))) return rc;
    if ((rc = fl->misc.ensure(256))) return rc;
    DeviceInfo info;
    if ((rc = device_info(&info))) return rc;
    // one workgroup per CU (the layer's weights and the operand slots fill
    // the LDS), each writing its own weight-gradient sums
    const int nparts = (int)std::min<uint64_t>((uint64_t)info.cus, (ntiles + kBwdWaves - 1) / kBwdWaves);
    if ((rc = fl->wpart.ensure((size_t)nparts * kPLayer * 4))) return rc;
    uint32_t* mx = (uint32_t*)fl->misc.ptr;
    hipLaunchKernelGGL(flow_zero_u32_kernel, dim3(1), dim3(64), 0, s, mx, 1);
    const float pf = loss == VGX_FLOW_LOSS_KL ? 1.f : 2.f, pq = loss == VGX_FLOW_LOSS_VARIANCE ? 1.f : 0.f;
    hipLaunchKernelGGL(flow_lossw_max_kernel, dim3(256), dim3(256), 0, s, w, f, (const float*)fl->logq.ptr,
                       (uint64_t)n, pf, pq, mx);
    hipLaunchKernelGGL(flow_lossw_sum_kernel, dim3(nc), dim3(256), 0, s, w, f, (const float*)fl->logq.ptr,
                       (uint64_t)n, pf, pq, (const uint32_t*)mx, (float*)fl->c.ptr, (double*)fl->cpart.ptr);
    hipLaunchKernelGGL(flow_csum_kernel, dim3(1), dim3(256), 0, s, (const double*)fl->cpart.ptr, nc,
                       (double*)fl->misc.ptr + 1);
    if ((rc = check_launch())) return rc;
    const size_t blds = lds_bwd();
    for (int k = 0; k < fl->L; ++k) {
        BwdArgs b{};
        b.n = (uint64_t)n;
        b.k = k;
        b.saved = (const float*)fl->saved.ptr;
        b.u0 = (const float*)fl->u0.ptr;
        b.c = (const float*)fl->c.ptr;
        b.csum = (const double*)fl->misc.ptr + 1;
        b.G = (float*)fl->G.ptr;
        b.part = (float*)fl->wpart.ptr;
        hipEvent_t e0, e1;
        if (timed_slot(g_timing_ctx, kTimedFlowBwd, &e0, &e1))
            hipExtLaunchKernelGGL(flow_backward_kernel, dim3(nparts), dim3(kBwdThreads), (uint32_t)blds, s, e0, e1, 0,
                                  flow_args(fl), b);
        else
            hipLaunchKernelGGL(flow_backward_kernel, dim3(nparts), dim3(kBwdThreads), blds, s, flow_args(fl), b);
        hipLaunchKernelGGL(flow_wgrad_reduce_kernel, dim3((kPLayer + kRedCols - 1) / kRedCols),
                           dim3(kRedCols * kRedGroups), 0, s,
                           (const float*)fl->wpart.ptr, nparts,
                           fl->grad + (size_t)k * kPLayer);
        if ((rc = check_launch())) return rc;
    }
    return VGX_OK;
}

extern "C" int vgx_flow_adam(vgx_flow* fl, float lr, float b1, float b2, float eps, float clip, void* stream) {
    if (!fl) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    fl->step += 1;
    AdamArgs a{};
    a.P = fl->P;
    a.grad = fl->grad;
    a.m = fl->m;
    a.v = fl->v;
    a.H = fl->H;
    a.HT = fl->HT;
    a.L = fl->L;
    a.lr = lr;
    a.b1 = b1;
    a.b2 = b2;
    a.eps = eps;
    a.c1 = 1.f - powf(b1, (float)fl->step);
    a.c2 = 1.f - powf(b2, (float)fl->step);
    a.clip = clip;
    hipLaunchKernelGGL(flow_adam_kernel, dim3(256), dim3(256), 0, s, a);
    int rc = check_launch();
    if (rc) return rc;
    return flow_cast(fl, s);
}

extern "C" int vgx_flow_eval(vgx_flow* fl, const vgx_integrand* fcn, const float* x, int64_t n, float* f,
                             void* stream) {
    if (!fl || !x || !f || n < 1) return VGX_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    FlowFcn F{};
    int rc = flow_fcn_params(fl, fcn, &F, s);
    if (rc) return rc;
    const size_t ncp = F.kind == VGX_FCN_DIAGRIDGE ? 3 * (size_t)F.ncomp : (size_t)F.ncomp * (2 + fl->D);
    const uint64_t ntiles = ((uint64_t)n + 15) / 16;
    const unsigned blocks = (unsigned)std::min<uint64_t>(4096, (ntiles + 3) / 4);
    if (hipFuncSetAttribute((const void*)flow_eval_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(ncp * 4 + 16)) != hipSuccess)
        return VGX_EHIP;
    hipLaunchKernelGGL(flow_eval_kernel, dim3(blocks), dim3(256), ncp * 4 + 16, s, F, fl->D, x, (uint64_t)n, f);
    return check_launch();
}

extern "C" int vgx_flow_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, void* stream) {
    if (!A || !B || !D) return VGX_EINVAL;
    hipLaunchKernelGGL(flow_mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D);
    return check_launch();
}
