// vgx_device.h -- device helpers shared by the libvgx kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vgx.h"

namespace vgx {

// Constants of the reference (src/vegas/_vegas.pyx:34-36).
constexpr double kTiny = 1e-257;
constexpr double kEpsilon = 2.220446049250313080847263336181640625e-16 * 1e4;
constexpr int kWave = 64;

// Training records of the fused sampler: one u16 bin per point and axis, or
// kNoBin where the point is not trained on the axis.
constexpr uint32_t kNoBin = 0xFFFFu;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter layout shared
// with the sampler: ctr = (p lo, p hi, iteration, block), key = (seed lo, hi).
struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// The same function for device code: each round's two three-input XORs
// (product half ^ counter word ^ key) are one v_bitop3_b32 each (truth
// table 0x96 = a ^ b ^ c), half the XOR instructions of the plain form.
__device__ __forceinline__ U4 philox4x32_10_dev(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 n;
        n.x = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c.y, k0, 0x96);
        n.y = (uint32_t)p1;
        n.z = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c.w, k1, 0x96);
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 24-bit uniform in [0,1) from one word (exact in float and double).
__host__ __device__ inline double u24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }
// 53-bit uniform in [0,1) from two words.
__host__ __device__ inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Division of 32-bit unsigned n by an invariant divisor d >= 1 with one
// multiply-high (Granlund & Montgomery, PLDI'94, section 4):
//   l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
//   q = (mulhi(m, n) + n) >> l      (33-bit intermediate; exact for all n).
struct DivMagic {
    uint32_t m, l, d, pad;
};

inline DivMagic make_divmagic(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    uint64_t m = (((1ull << 32) * ((1ull << l) - d)) / d) + 1;
    return DivMagic{(uint32_t)m, l, d, 0};
}

__host__ __device__ inline uint32_t divq(uint32_t n, const DivMagic& g) {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t hi = __umulhi(n, g.m);
#else
    uint32_t hi = (uint32_t)(((uint64_t)n * g.m) >> 32);
#endif
    return (uint32_t)(((uint64_t)hi + n) >> g.l);
}

// ---------------------------------------------------------------------------
// Compensated (double-double) accumulation: Knuth's TwoSum.  The library is
// built with -ffp-contract=off, so these are exact error-free transforms.
struct DD {
    double hi, lo;
};

__host__ __device__ inline void dd_add(DD& a, double b) {
    double s = a.hi + b;
    double bp = s - a.hi;
    double e = (a.hi - (s - bp)) + (b - bp);
    a.hi = s;
    a.lo += e;
}

__host__ __device__ inline DD dd_sum(DD a, DD b) {
    double s = a.hi + b.hi;
    double bp = s - a.hi;
    double e = (a.hi - (s - bp)) + (b.hi - bp);
    e += a.lo + b.lo;
    double hi = s + e;
    return DD{hi, e - (hi - s)};
}

__host__ __device__ inline double dd_value(DD a) { return a.hi + a.lo; }

// x^b for x >= 0, b > 0, from the FP32 hardware log2/exp2 (v_log_f32,
// v_exp_f32) with the binary exponent carried exactly: x = m 2^e with m in
// [0.5, 1), x^b = 2^(b (e + log2 m)).  Relative error below 3e-7 over the
// whole FP64 range; 0 for x = 0 (and for NaN).
__device__ inline double pow_fast(double x, double b) {
    if (!(x > 0.0)) return 0.0;
    if (x > 1.7976931348623157e308) return x;
    int e;
    const double m = frexp(x, &e);
    const double L = (double)e + (double)__builtin_amdgcn_logf((float)m);
    const double y = b * L;
    const double fl = floor(y);
    const float fr = (float)(y - fl);
    return ldexp((double)__builtin_amdgcn_exp2f(fr), (int)fl);
}

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64).
__device__ inline int lane_id() { return (int)__lane_id(); }

__device__ inline double wave_sum_fixed(double v) {
    // Butterfly in a fixed order: deterministic for a given lane assignment.
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ inline DD wave_sum_dd(DD v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        DD w{__shfl_xor(v.hi, o, kWave), __shfl_xor(v.lo, o, kWave)};
        // combine in lane-order-independent form: smaller lane index first
        v = (lane_id() & o) ? dd_sum(w, v) : dd_sum(v, w);
    }
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// ---- DPP lane movement (gfx9 encodings): another lane's value through the
// VALU's data-parallel-primitive path instead of an LDS permute.
constexpr int kDppRowShl = 0x100;     // + n: row_shl:n (within 16-lane rows)
constexpr int kDppRowShr = 0x110;     // + n: row_shr:n
constexpr int kDppWaveShl1 = 0x130;   // lane i reads lane i+1
constexpr int kDppWaveShr1 = 0x138;   // lane i reads lane i-1
constexpr int kDppRowBcast15 = 0x142; // lane 15 of row r to the lanes of row r+1
constexpr int kDppRowBcast31 = 0x143; // lane 31 to rows 2 and 3
// Lanes with no source (or outside ROWS) keep `old`.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_i(int old, int x) {
    return __builtin_amdgcn_update_dpp(old, x, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_d(double old, double x) {
    const uint64_t o = (uint64_t)__double_as_longlong(old), v = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)dpp_i<CTRL, ROWS>((int)(uint32_t)o, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)dpp_i<CTRL, ROWS>((int)(uint32_t)(o >> 32), (int)(uint32_t)(v >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// Another lane's value with every row enabled and 0 where the lane has no
// source (bound_ctrl): the value dpp_d(0.0, x) gives, without first writing
// the zero into the destination (an FP64 consumer cannot take the DPP
// modifier itself, so the old-value form costs two more moves per double).
template <int CTRL>
__device__ __forceinline__ int dpp0_i(int x) {
    return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp0_d(double x) {
    const uint64_t v = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)dpp0_i<CTRL>((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)dpp0_i<CTRL>((int)(uint32_t)(v >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// The FP64 sum of the 64 lanes' values in lane 63, through DPP moves (no
// LDS): quads (xor 1, xor 2), halves and rows (mirrors), then rows 1 and 3
// take rows 0 and 2 (lane 15 broadcast) and rows 2 and 3 take row 1 (lane
// 31).  FP64 addition is commutative, so the lanes of one row agree and the
// result is a fixed function of the 64 inputs (a fixed summation tree).
__device__ inline double wave_total(double v) {
    v += dpp0_d<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp0_d<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp0_d<0x141>(v);   // row_half_mirror
    v += dpp0_d<0x140>(v);   // row_mirror
    v += dpp_d<kDppRowBcast15, 0xA>(0.0, v);
    v += dpp_d<kDppRowBcast31, 0xC>(0.0, v);
    return v;
}

// dpp_d(0.0, x) in its cheaper form where every row is enabled.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_z(double x) {
    if constexpr (ROWS == 0xF) return dpp0_d<CTRL>(x);
    else return dpp_d<CTRL, ROWS>(0.0, x);
}

// One step of a segmented inclusive scan (v += v of the source lane unless a
// segment starts at or before this lane since the source), for 1 or 2 values.
template <int CTRL, int ROWS>
__device__ __forceinline__ void segscan_step(double& v, int& f) {
    const double pv = dpp_z<CTRL, ROWS>(v);
    const int pf = dpp_i<CTRL, ROWS>(0, f);
    if (!f) v = pv + v;
    f |= pf;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void segscan_step2(double& a, double& b, int& f) {
    const double pa = dpp_z<CTRL, ROWS>(a), pb = dpp_z<CTRL, ROWS>(b);
    const int pf = dpp_i<CTRL, ROWS>(0, f);
    if (!f) {
        a = pa + a;
        b = pb + b;
    }
    f |= pf;
}

// Segmented inclusive scan across the 64 lanes.  head = 1 when the lane
// starts a new segment.  Fixed order of additions: within each 16-lane row
// by shifts of 1, 2, 4, 8, then the carry of row 0 into row 1 and of row 2
// into row 3, then that of rows 0-1 into rows 2-3.
__device__ inline double wave_segscan(double v, int head) {
    int f = head;
    segscan_step<kDppRowShr + 1, 0xF>(v, f);
    segscan_step<kDppRowShr + 2, 0xF>(v, f);
    segscan_step<kDppRowShr + 4, 0xF>(v, f);
    segscan_step<kDppRowShr + 8, 0xF>(v, f);
    segscan_step<kDppRowBcast15, 0xA>(v, f);
    segscan_step<kDppRowBcast31, 0xC>(v, f);
    return v;
}
__device__ inline void wave_segscan2(double& a, double& b, int head) {
    int f = head;
    segscan_step2<kDppRowShr + 1, 0xF>(a, b, f);
    segscan_step2<kDppRowShr + 2, 0xF>(a, b, f);
    segscan_step2<kDppRowShr + 4, 0xF>(a, b, f);
    segscan_step2<kDppRowShr + 8, 0xF>(a, b, f);
    segscan_step2<kDppRowBcast15, 0xA>(a, b, f);
    segscan_step2<kDppRowBcast31, 0xC>(a, b, f);
}

}  // namespace vgx
