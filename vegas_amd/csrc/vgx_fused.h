// vgx_fused.h -- the fused vegas+ sampler kernel (template; instantiated per
// dimension and integrand in vgx_fused_*.hip).
//
// One wave owns a tile of kT consecutive points of the iteration (kK per
// lane, lane-contiguous).  Per point: Philox4x32-10 uniforms -> stratified y
// (pyx:1734-1741) -> map with the grid in LDS (pyx:347-359) -> weight
// J*(dv_y/n_h) (pyx:1749) -> integrand -> w f.  Per hcube: the reference's
// two-pass sums with the EPSILON guard (pyx:2146-2183) as wave-level
// segmented scans over the tile; hcubes cut by a tile edge are finished by
// the straddler kernel from per-tile part sums.  Training data (w f n)^2 (pyx:2173) is written
// as one record per point (its bin on every axis and the value), from which a separate pass
// builds the histogram in a fixed order; the batched sampler, which calls the same tile
// body, keeps an LDS-privatised histogram (FP64 sums, packed u16 counts) instead.
#pragma once
#include "vgx_iter.h"

namespace vgx {

// Scalars of one integral's iteration that the sampler and the tile body
// read (a kernel-argument IterGeom, or one integral of a batch).
struct TileGeom {
    int dim, ld, nb;
    int exact_div, hist_lds, train_points, adaptive_strat, fast_pow, train_errors;
    uint32_t iteration, seed_lo, seed_hi;
    double beta_half, dv_y;
    double jn;  // product of ninc over the axes (the integer stratum path's Jacobian factor)
    const double* uext;  // caller's uniforms [total][dim] for the 53-bit path, or null
};

__host__ __device__ inline TileGeom tile_geom(const IterGeom& g) {
    TileGeom t;
    t.dim = g.dim;
    t.ld = g.ld;
    t.nb = g.nb;
    t.exact_div = g.exact_div;
    t.hist_lds = g.hist_lds;
    t.train_points = g.train_points;
    t.adaptive_strat = g.adaptive_strat;
    t.fast_pow = g.fast_pow;
    t.train_errors = g.train_errors;
    t.iteration = g.iteration;
    t.seed_lo = g.seed_lo;
    t.seed_hi = g.seed_hi;
    t.beta_half = g.beta_half;
    t.dv_y = g.dv_y;
    t.jn = g.jn;
    t.uext = g.uext;
    return t;
}

// Per-axis constants of the sampler (device memory, written before the
// sampler runs and read with scalar loads).
struct AxisC {
    uint32_t m, l, nstrat, ninc;  // DivMagic of nstrat, strata and increments
    double inv_s;                 // 1 / nstrat
    uint32_t per;                 // ninc / nstrat when ninc is a multiple of nstrat, else 0
    uint32_t pad;
};

__device__ __forceinline__ AxisC ldc(const AxisC* p, int i) {
    const uint32_t* u = reinterpret_cast<const uint32_t*>(p + i);
    const double* f = reinterpret_cast<const double*>(p + i);
    return AxisC{ldc(u, 0), ldc(u, 1), ldc(u, 2), ldc(u, 3), ldc(f, 2), ldc(u, 6), 0u};
}

// The LDS counts hold bins 2i and 2i+1 as the low and high halves of word i:
// bin idx adds 1 << (16 (idx & 1)).  (idx << 4) & 31 is that shift, in the
// form the 32-bit shift takes without a mask (it reads 5 bits of the count).
__device__ __forceinline__ uint32_t pair_inc(int idx) { return 1u << (((uint32_t)idx << 4) & 31u); }

// Stratum coordinates of hcube h, axis 0 fastest (pyx:1736-1738), each
// premultiplied by the axis's increments per stratum (the first increment
// index of the stratum, < 2^16).
template <int DIM>
__device__ __forceinline__ void per_coords(const AxisC* __restrict__ ax, int dim, uint32_t h,
                                           uint32_t* cc) {
    uint32_t t = h;
#pragma unroll
    for (int d = 0; d < nd_max(DIM); ++d) {
        if (DIM <= 0 && d >= dim) break;
        const AxisC a = ldc(ax, d);
        const uint32_t q = (uint32_t)(((uint64_t)__umulhi(t, a.m) + t) >> a.l);
        const uint32_t c = t - q * a.nstrat;
        t = q;
        cc[d] = c * a.per;
    }
}

// The PER coordinates of one hcube, two 16-bit fields per register: a lane
// computes them once per hcube and reuses them for its points in that hcube.
template <int DIM>
struct PackedCoords {
    static constexpr int kN = (nd_max(DIM) + 1) / 2;
    uint32_t v[kN];
    __device__ __forceinline__ void set(const AxisC* __restrict__ ax, int dim, uint32_t h) {
        uint32_t cc[2 * kN];
#pragma unroll
        for (int d = 0; d < 2 * kN; ++d) cc[d] = 0u;
        per_coords<DIM>(ax, dim, h, cc);
#pragma unroll
        for (int i = 0; i < kN; ++i) v[i] = cc[2 * i] | (cc[2 * i + 1] << 16);
    }
    __device__ __forceinline__ uint32_t get(int d) const { return (v[d >> 1] >> (16 * (d & 1))) & 0xFFFFu; }
    // h -> h + 1: axis 0 advances and carries into the next axes (the digit
    // order of pyx:1736-1738), with no division.
    __device__ __forceinline__ void step(const AxisC* __restrict__ ax, int dim) {
#pragma unroll
        for (int d = 0; d < nd_max(DIM); ++d) {
            if (DIM <= 0 && d >= dim) break;
            const AxisC a = ldc(ax, d);
            const uint32_t c = get(d) + a.per;
            const uint32_t sh = 16 * (d & 1);
            const bool carry = c >= a.ninc;
            v[d >> 1] = (v[d >> 1] & ~(0xFFFFu << sh)) | ((carry ? 0u : c) << sh);
            if (!carry) break;
        }
    }
};

// Coordinates of hcube h decoded axis by axis as eval_point walks the axes
// (no array of them stays live).
struct CoordWalk {
    uint32_t t;
    __device__ __forceinline__ uint32_t next(const AxisC& a) {
        const uint32_t q = (uint32_t)(((uint64_t)__umulhi(t, a.m) + t) >> a.l);
        const uint32_t c = t - q * a.nstrat;
        t = q;
        return c;
    }
};

// Generates uniforms, y, x and J for point p of hcube h (with PER, pc holds
// h's premultiplied coordinates).
//   PER (24-bit uniforms, every axis with ninc = per * nstrat, no exact
//       division): the increment index and offset follow exactly from
//       t = (c + u) * per with 24-bit integer multiplies:
//       k = c*per + floor(u24*per / 2^24), offset = the low 24 bits * 2^-24
//       (the map of the exact y = (c+u)/nstrat); k < ninc always;
//   otherwise (53-bit uniforms, or unaligned axes, or exact_div) y is formed
//       in FP64 as the reference does, y = (c+u)/nstrat (pyx:1741), and the
//       map computes t = y*ninc (pyx:351); the 53-bit uniforms are the
//       caller's row p of g.uext when it is set (ran_array_generator,
//       pyx:1732), else Philox4x32-10 words;
//   k = ninc (y = 1) takes x = grid[ninc] and the last increment
//   (pyx:355-359).
//   kbin[d]: the training bin of axis d, -1 where y is not inside (0,1)
//   (pyx:459-463); with PER the bins are always given and zmin (min over the
//   axes of the 24-bit uniform, to be initialised by the caller) is 0 when
//   some axis may have y = 0 exactly (y = 0 needs u = 0 and stratum 0) --
//   per_zero_axes() then names the axes to skip, if any.
//   perv (PER, optional): the axes' increments per stratum, held in vector
//   registers by the caller so that no scalar load per point and axis sits
//   between the grid reads (a scalar load shares the LDS reads' wait counter
//   and returns out of order, which would wait out every grid read at once).
//   DIM > 0: that dimension; DIM <= 0: runtime g.dim, at most nd_max(DIM).
template <int DIM, bool R32, bool PER>
__device__ __forceinline__ void eval_point(const TileGeom& g, const AxisC* __restrict__ ax,
                                           const double* __restrict__ G, const PackedCoords<DIM>& pc,
                                           uint32_t h, uint32_t p, double* x, double& J, int* kbin,
                                           uint32_t& zmin, const uint32_t* perv = nullptr) {
    static_assert(R32 || !PER, "PER needs 24-bit uniforms");
    const int dim = DIM > 0 ? DIM : g.dim;
    U4 w{0, 0, 0, 0};
    J = 1.0;
    CoordWalk walk{h};
#pragma unroll
    for (int d = 0; d < nd_max(DIM); ++d) {
        if (DIM <= 0 && d >= dim) break;
        const AxisC a = ldc(ax, d);
        const uint32_t c = PER ? pc.get(d) : walk.next(a);
        const int ni = (int)a.ninc;
        int k, kv;
        double dl;
        if (R32) {
            if ((d & 3) == 0) w = philox4x32_10_dev(U4{p, 0u, g.iteration, (uint32_t)(d >> 2)}, g.seed_lo, g.seed_hi);
            const int e = d & 3;
            const uint32_t wd = e == 0 ? w.x : e == 1 ? w.y : e == 2 ? w.z : w.w;
            const uint32_t ui = wd >> 8;
            if (PER) {
                // both factors < 2^24: the backend selects the 24-bit multiplies
                const uint32_t pv = perv ? perv[d] : (a.per & 0xFFFFFu);
                const uint64_t prod = (uint64_t)ui * (uint64_t)pv;
                k = (int)(c + (uint32_t)(prod >> 24));
                dl = (double)((uint32_t)prod & 0xFFFFFFu) * (1.0 / 16777216.0);
                kv = k;  // y = 0 (c = 0, u = 0) is screened through zmin
                zmin = min(zmin, ui);
            } else {
                const double yv = g.exact_div ? ((double)c + u24(wd)) / (double)a.nstrat
                                              : ((double)c + u24(wd)) * a.inv_s;
                const double tt = fmin(yv * (double)ni, (double)ni);
                k = (int)tt;
                dl = tt - (double)k;
                kv = (yv > 0 && yv < 1) ? (k < ni ? k : ni - 1) : -1;
            }
        } else {
            double u;
            if (g.uext) {
                u = g.uext[(size_t)p * dim + d];
            } else {
                if ((d & 1) == 0)
                    w = philox4x32_10_dev(U4{p, 0u, g.iteration, (uint32_t)(d >> 1)}, g.seed_lo, g.seed_hi);
                u = (d & 1) ? u53(w.z, w.w) : u53(w.x, w.y);
            }
            const double yv = g.exact_div ? ((double)c + u) / (double)a.nstrat : ((double)c + u) * a.inv_s;
            const double tt = fmin(yv * (double)ni, (double)ni);
            k = (int)tt;
            dl = tt - (double)k;
            kv = (yv > 0 && yv < 1) ? (k < ni ? k : ni - 1) : -1;
        }
        const double* gd = G + d * g.ld;
        if (PER) {
            // k < ninc always: no branch, the loads can be issued early; x by
            // one FMA, and the ninc factors of J applied once (g.jn) after
            // the axes (the FP64 parity mode takes the branch below)
            const double g0 = gd[k], g1 = gd[k + 1];
            const double inc = g1 - g0;
            x[d] = __builtin_fma(inc, dl, g0);
            J *= inc;
        } else if (k < ni) {
            const double g0 = gd[k], g1 = gd[k + 1];
            const double inc = g1 - g0;
            x[d] = g0 + inc * dl;
            J *= inc * (double)ni;
        } else {
            x[d] = gd[ni];
            J *= (gd[ni] - gd[ni - 1]) * (double)ni;
        }
        kbin[d] = kv;
    }
    if (PER) J *= g.jn;
}
// PER: bit d set where point p of the hcube with coordinates pc has y = 0
// exactly on axis d (stratum 0 and a zero uniform; probability 2^-24 per
// axis and point, so this regenerates the uniforms instead of carrying a
// mask through the point loop).
template <int DIM>
__device__ __forceinline__ uint64_t per_zero_axes(const TileGeom& g, const PackedCoords<DIM>& pc, uint32_t p) {
    const int dim = DIM > 0 ? DIM : g.dim;
    uint64_t m = 0;
    U4 w{0, 0, 0, 0};
    for (int d = 0; d < nd_max(DIM); ++d) {
        if (DIM <= 0 && d >= dim) break;
        if ((d & 3) == 0) w = philox4x32_10_dev(U4{p, 0u, g.iteration, (uint32_t)(d >> 2)}, g.seed_lo, g.seed_hi);
        const int e = d & 3;
        const uint32_t wd = e == 0 ? w.x : e == 1 ? w.y : e == 2 ? w.z : w.w;
        if ((pc.get(d) | (wd >> 8)) == 0) m |= 1ull << d;
    }
    return m;
}

// y of point p of hcube h, formed as eval_point forms it (staged sampler).
template <bool R32>
__device__ inline void point_y(const TileGeom& g, const AxisC* __restrict__ ax, uint32_t h,
                               uint32_t p, double* y) {
    uint32_t t = h;
    U4 w{0, 0, 0, 0};
    for (int d = 0; d < g.dim; ++d) {
        const AxisC a = ax[d];
        const uint32_t q = (uint32_t)(((uint64_t)__umulhi(t, a.m) + t) >> a.l);
        const uint32_t c = t - q * a.nstrat;
        t = q;
        double u;
        if (R32) {
            if ((d & 3) == 0) w = philox4x32_10_dev(U4{p, 0u, g.iteration, (uint32_t)(d >> 2)}, g.seed_lo, g.seed_hi);
            const int e = d & 3;
            u = u24(e == 0 ? w.x : e == 1 ? w.y : e == 2 ? w.z : w.w);
        } else if (g.uext) {
            u = g.uext[(size_t)p * g.dim + d];
        } else {
            if ((d & 1) == 0) w = philox4x32_10_dev(U4{p, 0u, g.iteration, (uint32_t)(d >> 1)}, g.seed_lo, g.seed_hi);
            u = (d & 1) ? u53(w.z, w.w) : u53(w.x, w.y);
        }
        y[d] = g.exact_div ? ((double)c + u) / (double)a.nstrat : ((double)c + u) * a.inv_s;
    }
}

__host__ __device__ inline AxisC make_axis(const IterGeom& g, int d) {
    const uint32_t s = (uint32_t)g.nstrat[d], n = (uint32_t)g.ninc[d];
    const uint32_t per = (n % s == 0 && n / s < (1u << 20)) ? n / s : 0u;
    return AxisC{g.dm[d].m, g.dm[d].l, s, n, g.inv_nstrat[d], per, 0u};
}

// Two-value version of seg_prefix (shares the segment flags).
__device__ __forceinline__ void seg_prefix2(const int (&key)[kK], double (&a)[kK], double (&b)[kK],
                                            int lane) {
#pragma unroll
    for (int j = 1; j < kK; ++j)
        if (key[j] == key[j - 1]) {
            a[j] += a[j - 1];
            b[j] += b[j - 1];
        }
    const int prev_last = dpp_i<kDppWaveShr1>(-1, key[kK - 1]);  // lane 0: -1
    double va = a[kK - 1], vb = b[kK - 1];
    wave_segscan2(va, vb, key[kK - 1] != prev_last);
    const double ca = dpp0_d<kDppWaveShr1>(va), cb = dpp0_d<kDppWaveShr1>(vb);
    const bool carry = lane > 0 && key[0] == prev_last;
#pragma unroll
    for (int j = 0; j < kK; ++j)
        if (carry && key[j] == key[0]) {
            a[j] += ca;
            b[j] += cb;
        }
}

// Sum of v over the segment of equal keys up to and including each element
// (lane-contiguous elements, keys non-decreasing across the wave).
__device__ __forceinline__ void seg_prefix(const int (&key)[kK], const double (&v)[kK],
                                           double (&P)[kK], int lane) {
    P[0] = v[0];
#pragma unroll
    for (int j = 1; j < kK; ++j) P[j] = (key[j] == key[j - 1]) ? P[j - 1] + v[j] : v[j];
    const int prev_last = dpp_i<kDppWaveShr1>(-1, key[kK - 1]);  // lane 0: -1
    const int head = key[kK - 1] != prev_last;
    const double incl = wave_segscan(P[kK - 1], head);
    const double prev_incl = dpp0_d<kDppWaveShr1>(incl);
    const double cin = (lane > 0 && key[0] == prev_last) ? prev_incl : 0.0;
#pragma unroll
    for (int j = 0; j < kK; ++j)
        if (key[j] == key[0]) P[j] += cin;
}

// Global pointers the tile body writes or reads.
struct TileIO {
    const uint32_t* off;     // point offset of every hcube
    double* sigf;
    Slot* slots;
    double* errv;
    uint32_t* rec_bins;      // training records (single integral), or null: LDS histogram
    double* rec_c2;
    uint32_t rec_p0;         // point index of record 0
    uint32_t rec_stride;     // records per axis array
};

// The training record of point ri: its bins on axes 2k and 2k+1 as the low
// and high halves of rec_bins[k * rec_stride + ri] (kNoBin where kb[d] < 0,
// where bit d of zm is set, and past the last axis), so that a pass over one
// axis reads one contiguous array; c2 = (w f n)^2 at rec_c2[ri].
template <int DIM>
__device__ __forceinline__ void store_record(const TileIO& io, int dim, uint32_t ri, const int* kb,
                                             uint64_t zm, double c2) {
    constexpr int ND = nd_max(DIM);
    constexpr int NW = (ND + 1) / 2;
    uint32_t wd[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t lo = 2 * k < dim ? (uint32_t)kb[2 * k] : kNoBin;
        const uint32_t hi = (2 * k + 1 < ND && 2 * k + 1 < dim) ? (uint32_t)kb[2 * k + 1 < ND ? 2 * k + 1 : 0] : kNoBin;
        // low 16 bits of lo, then of hi (one v_perm_b32)
        wd[k] = __builtin_amdgcn_perm(hi, lo, 0x05040100u);
    }
    if (zm) {  // y = 0 exactly on some axis (rare): those axes are not trained on
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            if ((zm >> (2 * k)) & 1) wd[k] |= kNoBin;
            if ((zm >> (2 * k + 1)) & 1) wd[k] |= kNoBin << 16;
        }
    }
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        if (DIM <= 0 && 2 * k >= dim) break;
        // a uniform row base and a 32-bit lane offset (no 64-bit address math per lane)
        uint32_t* row = io.rec_bins + (size_t)k * io.rec_stride;
        row[ri] = wd[k];
    }
    io.rec_c2[ri] = c2;
}

// Per-wave accumulators across tiles.
struct TileAcc {
    DD am, av, as;           // sum of S + D, of var_h, of sigf_h
    double wmax;
    int nan_seen;
};

// Where the tile sits: global tile t with points [plo, phi), first hcube hf;
// hcubes [.., hend) of the batch (offset off_end at hend); the integral's
// first hcube hbase and the shift from global to integral-local point index
// (both 0 for a single integral).
struct TilePos {
    uint32_t t, plo, phi, hf, hend, off_end, hbase, pshift;
};

// Offsets of the hcubes from hf on, as a wave tile needs them: entry i of
// lane i + kWave*q (past hend: the end offset).
constexpr int kOffRegs = (kNoff + kWave - 1) / kWave;
__device__ __forceinline__ void load_offsets(const uint32_t* __restrict__ off, uint32_t hf, uint32_t hend,
                                             uint32_t off_end, int lane, uint32_t (&v)[kOffRegs]) {
#pragma unroll
    for (int q = 0; q < kOffRegs; ++q) {
        const uint32_t hh = hf + (uint32_t)(lane + kWave * q);
        v[q] = hh < hend ? off[hh] : off_end;
    }
}
__device__ __forceinline__ void store_offsets(uint32_t* offs, int lane, const uint32_t (&v)[kOffRegs]) {
#pragma unroll
    for (int q = 0; q < kOffRegs; ++q)
        if (lane + kWave * q < kNoff) offs[lane + kWave * q] = v[q];
}

// One wave tile (kT points, kK per lane): sample, map, integrand, training
// histogram, and the reference's two passes per hcube as segmented scans.
template <int DIM, bool PER, class Fcn>
__device__ __forceinline__ void tile_body(const TileGeom& g, const TileIO& io, const Fcn& fcn,
                                          const AxisC* __restrict__ ax, const double* __restrict__ G,
                                          double* Ls, uint32_t* Lc, uint32_t* offs, const TilePos& tp,
                                          int lane, TileAcc& acc) {
    constexpr bool R32 = Fcn::kRng32;
    constexpr int ND = nd_max(DIM);
    const int dim = DIM > 0 ? DIM : g.dim;
    const double eps = kEpsilon;
    const uint32_t t = tp.t, plo = tp.plo, phi = tp.phi, hf = tp.hf;
    // offs[] (the offsets of hcubes hf, hf+1, ...) is staged by the caller
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // hcube (local index) of this lane's first point
    const uint32_t pl0 = plo + lane * kK;
    int li = 0;
    {
        int lo = 0, hi = kNoff - 2;  // largest i with offs[i] <= pl0
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= pl0) lo = mid; else hi = mid - 1;
        }
        li = lo;
    }
    // Points are evaluated one at a time (no unrolling: one point's
    // registers live at once); results rotate through a register window so
    // element j ends in key[j], wf[j], nn[j].
    int key[kK];
    double wf[kK];
    uint32_t nn[kK];
#pragma unroll
    for (int j = 0; j < kK; ++j) {
        key[j] = kKeySentinel;
        wf[j] = 0.0;
        nn[j] = 1;
    }
    // hcube cur_i (local index) of the lane's current point: its end offset
    // nxt, its point count, dv_y / n_h (pyx:1749) and, PER, its coordinates
    int cur_i = li;
    uint32_t nxt = 0xFFFFFFFFu;
    double wdiv = 0.0;
    uint32_t ncur = 1;
    PackedCoords<DIM> pc;
    if (pl0 < phi) {
        nxt = offs[li + 1];
        ncur = nxt - offs[li];
        wdiv = g.dv_y / (double)ncur;
        if (PER) pc.set(ax, dim, hf + li - tp.hbase);
    }
    // PER with the dimension known (up to 8 axes; above, the registers would
    // spill): each axis's increments per stratum in a vector register for
    // the tile (the empty asm leaves it there)
    constexpr bool kPinPer = PER && DIM > 0 && DIM <= 8;
    uint32_t perv[ND];
    if constexpr (kPinPer) {
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            uint32_t v = ldc(ax, d).per;
            asm volatile("" : "+v"(v));
            perv[d] = v & 0xFFFFFu;
        }
    }
#pragma unroll 1
    for (int j = 0; j < kK; ++j) {
        const uint32_t pj = pl0 + j;
        int kj = kKeySentinel;
        double vj = 0.0;
        uint32_t nj = 1;
        if (pj < phi) {
            while (pj >= nxt && cur_i < kNoff - 2) {  // next hcube (one step: hcubes have >= 2 points)
                ++cur_i;
                const uint32_t o0 = nxt;
                nxt = offs[cur_i + 1];
                ncur = nxt - o0;
                wdiv = g.dv_y / (double)ncur;
                if (PER) pc.step(ax, dim);
            }
            kj = cur_i;
            nj = ncur;
            double x[ND];
            int kb[ND];
            double J;
            uint32_t zmin = 0xFFFFFFFFu;
            eval_point<DIM, R32, PER>(g, ax, G, pc, hf + cur_i - tp.hbase, pj - tp.pshift, x, J, kb, zmin,
                                      kPinPer ? perv : nullptr);
            const double w = J * wdiv;
            const double f = fcn.template eval<DIM>(x, dim, pj);
            if (f != f) acc.nan_seen = 1;
            vj = w * f;
            const double tn = vj * (double)ncur;
            acc.wmax = fmax(acc.wmax, fabs(tn));
            if (g.train_points) {
                const double c2 = tn * tn;
                if (io.rec_c2) {
                    // a record per point; the histogram sums are built from
                    // them in a fixed order (here, in tile order, or after the
                    // launch); the counts, integers, need no order and go
                    // into the LDS at once when Lc is given
                    const uint64_t zm = (PER && zmin == 0) ? per_zero_axes<DIM>(g, pc, pj - tp.pshift) : 0ull;
                    store_record<DIM>(io, dim, pj - io.rec_p0, kb, zm, c2);
                    if (Lc) {
                        if (PER && zm == 0) {
                            // every axis inside (0,1): no per-axis test
#pragma unroll
                            for (int d = 0; d < ND; ++d) {
                                if (DIM <= 0 && d >= dim) break;
                                const int idx = d * g.nb + kb[d];
                                atomicAdd(&Lc[idx >> 1], pair_inc(idx));
                            }
                        } else {
#pragma unroll
                            for (int d = 0; d < ND; ++d) {
                                if (DIM <= 0 && d >= dim) break;
                                const int k = kb[d];
                                if (k < 0 || ((zm >> d) & 1)) continue;
                                const int idx = d * g.nb + k;
                                atomicAdd(&Lc[idx >> 1], pair_inc(idx));
                            }
                        }
                    }
                } else if (PER && zmin != 0) {
                    // LDS histogram, every axis inside (0,1): no per-axis test
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        if (DIM <= 0 && d >= dim) break;
                        const int idx = d * g.nb + kb[d];
                        unsafeAtomicAdd(&Ls[idx], c2);
                        atomicAdd(&Lc[idx >> 1], pair_inc(idx));
                    }
                } else {
                    const uint64_t zm = (PER && zmin == 0) ? per_zero_axes<DIM>(g, pc, pj - tp.pshift) : 0ull;
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        if (DIM <= 0 && d >= dim) break;
                        const int k = kb[d];
                        if (k < 0 || ((zm >> d) & 1)) continue;
                        const int idx = d * g.nb + k;
                        unsafeAtomicAdd(&Ls[idx], c2);
                        atomicAdd(&Lc[idx >> 1], pair_inc(idx));
                    }
                }
            }
        }
        // rotate: element j enters at the back of the window
#pragma unroll
        for (int q = 0; q < kK - 1; ++q) {
            key[q] = key[q + 1];
            wf[q] = wf[q + 1];
            nn[q] = nn[q + 1];
        }
        key[kK - 1] = kj;
        wf[kK - 1] = vj;
        nn[kK - 1] = nj;
    }
    // ---- pass 1: S per segment (pyx:2147-2152)
    double P[kK];
    seg_prefix(key, wf, P, lane);
    int endf[kK];
    const int next_first = dpp_i<kDppWaveShl1>(kKeySentinel, key[0]);  // lane 63: unused
#pragma unroll
    for (int j = 0; j < kK - 1; ++j) endf[j] = key[j + 1] != key[j];
    endf[kK - 1] = (lane == kWave - 1) || (next_first != key[kK - 1]);
    // Total of the lane's last segment.  If it ends in this lane it is
    // P[kK-1]; otherwise it ends inside the first run of the nearest later
    // lane whose first run ends there: copy that run's closing prefix
    // backwards across the lanes that only continue.
    double fv = P[0];
#pragma unroll
    for (int j = 1; j < kK; ++j)
        if (key[j] == key[0]) fv = P[j];
    double rv = fv;
    int rf = (key[0] != key[kK - 1]) || endf[kK - 1];
    // Within each 16-lane row by DPP (a lane with no source reads value and
    // flag 0; a value is only used once its flag is set, so that is as if it
    // kept its own): lane i then holds the first flagged lane of [i, end of
    // its row].  Across rows by permutes from the first lane of the row 1,
    // then 2 rows ahead, which extends that to [i, 63].
    auto back_step = [&](double nv, int nf) {
        if (!rf) {
            rv = nv;
            rf = nf;
        }
    };
    back_step(dpp0_d<kDppRowShl + 1>(rv), dpp0_i<kDppRowShl + 1>(rf));
    back_step(dpp0_d<kDppRowShl + 2>(rv), dpp0_i<kDppRowShl + 2>(rf));
    back_step(dpp0_d<kDppRowShl + 4>(rv), dpp0_i<kDppRowShl + 4>(rf));
    back_step(dpp0_d<kDppRowShl + 8>(rv), dpp0_i<kDppRowShl + 8>(rf));
#pragma unroll
    for (int hop = 1; hop <= 2; hop <<= 1) {
        const int src = ((lane >> 4) + hop) << 4;
        const double nv = __shfl(rv, src & (kWave - 1), kWave);
        const int nf = __shfl(rf, src & (kWave - 1), kWave);
        if (src < kWave) back_step(nv, nf);
    }
    const double next_rv = dpp0_d<kDppWaveShl1>(rv);
    const double vlast = endf[kK - 1] ? P[kK - 1] : next_rv;
    double S[kK];
    {
        double c = vlast;
#pragma unroll
        for (int j = kK - 1; j >= 0; --j) {
            if (j < kK - 1 && endf[j]) c = P[j];
            S[j] = c;
        }
    }
    // straddling segments: first (local 0) and last (kL)
    int kL;
    {
        int lo = 0, hi = kNoff - 2;  // largest i with offs[i] < phi
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (offs[mid] < phi) lo = mid; else hi = mid - 1;
        }
        kL = lo;
    }
    const bool sA = offs[0] < plo || offs[1] > phi;
    const bool sB = offs[kL + 1] > phi;
    // ---- pass 2: guarded deviations (pyx:2153-2169).  A cut hcube's part
    // uses the part's own mean; the straddler kernel combines the parts.
    double qv[kK], dv[kK];
#pragma unroll
    for (int j = 0; j < kK; ++j) {
        qv[j] = 0.0;
        dv[j] = 0.0;
        const int k = key[j];
        if (k == kKeySentinel) continue;
        const bool strad = (k == 0 && sA) || (k == kL && sB);
        const uint32_t cnt = strad ? min(offs[k + 1], phi) - max(offs[k], plo) : nn[j];
        const double m = S[j] / (double)cnt;
        double dd = wf[j] - m;
        const double thr = eps * fabs(m);
        if (fabs(dd) < thr) {
            qv[j] = thr * thr;
            dv[j] = 0.0;
        } else {
            qv[j] = dd * dd;
            dv[j] = dd;
        }
    }
    seg_prefix2(key, qv, dv, lane);
    // Segment ends, one element at a time: element j sits at index 0 of the
    // rotating windows.
#pragma unroll 1
    for (int j = 0; j < kK; ++j) {
        const int k = key[0];
        if (endf[0] && k != kKeySentinel) {
            const bool strad = (k == 0 && sA) || (k == kL && sB);
            const uint32_t h = hf + k;
            if (!strad) {
                const double n = (double)nn[0];
                const double Q = qv[0], D = dv[0];
                const double varc = (n * Q - D * D) / (n - 1.);
                dd_add(acc.am, S[0] + D);
                dd_add(acc.av, varc);
                // sigf_h = |var_h|^(beta/2) (pyx:2183-2186); in exact mode a
                // separate FP64 pass raises it to the power
                if (g.adaptive_strat) {
                    if (g.fast_pow) {
                        const double sv = pow_fast(fabs(varc), g.beta_half);
                        io.sigf[h] = sv;
                        dd_add(acc.as, sv);
                    } else {
                        io.sigf[h] = fabs(varc);
                    }
                }
                if (g.train_errors) io.errv[h] = fabs(varc);
            }
            // cut hcubes: the tile's first segment and its last one
            const bool slA = k == 0, slB = pl0 + j == phi - 1;
            if (slA || slB) {
                Slot sl{-1, 0u, 0.0, 0.0, 0.0};
                if (strad && (slA ? sA : kL != 0)) {
                    sl.h = (int32_t)h;
                    sl.n = min(offs[k + 1], phi) - max(offs[k], plo);
                    sl.S = S[0];
                    sl.Q = qv[0];
                    sl.D = dv[0];
                }
                io.slots[2 * (size_t)t + (slA ? 0 : 1)] = sl;
                if (slA && slB) io.slots[2 * (size_t)t + 1].h = -1;
            }
        }
#pragma unroll
        for (int q = 0; q < kK - 1; ++q) {
            key[q] = key[q + 1];
            endf[q] = endf[q + 1];
            nn[q] = nn[q + 1];
            S[q] = S[q + 1];
            qv[q] = qv[q + 1];
            dv[q] = dv[q + 1];
        }
    }
    __builtin_amdgcn_wave_barrier();
}


// The u16 pair-packed LDS counts into a workgroup's own row of u32 counts
// (written, or added to when add), one lane per pair.
__device__ inline void counts_to_row(uint32_t* Lc, int tot, uint32_t* row, bool add, bool clear) {
    for (int i = threadIdx.x; i < (tot + 1) / 2; i += blockDim.x) {
        const uint32_t wv = Lc[i];
        const uint32_t lo = wv & 0xffffu, hi = wv >> 16;
        if (add) {
            row[2 * i] += lo;
            if (2 * i + 1 < tot) row[2 * i + 1] += hi;
        } else {
            row[2 * i] = lo;
            if (2 * i + 1 < tot) row[2 * i + 1] = hi;
        }
        if (clear) Lc[i] = 0u;
    }
}

// ---- training histogram in a fixed order ----------------------------------
// The waves of a workgroup take its tiles n = 0, 1, ... in turn from an LDS
// counter, so faster waves take more of them.  The axes are taken in pairs
// (the two halves of one record word); each pair k has its own adder lock
// and its own next tile.  The tiles add their (w f n)^2 to pair k's LDS sums
// strictly in tile order, each tile axis by axis and lane slot by lane slot
// (the LDS resolves the lanes of one instruction that hit one bin in a
// fixed order), so no bin's sum depends on timing.  (The counts are
// integers: the tile body adds them at once.)  No wave waits for another
// while it computes: a wave that has stored a tile's records marks the tile
// complete in a ring of kRing flags; then, for each pair in turn (starting
// at a pair of its own), if the pair's next tile is complete and no other
// wave holds the pair's lock, it adds that tile and the complete ones after
// it (up to kHoldTiles) under the lock, kBatch tiles' records loaded at
// once (the first batch before the lock is taken); different pairs are
// added by different waves at once.  A wave about to take tile n >= kRing
// first makes sure every pair has added tile n - kRing (whose flag slot it
// reuses), adding meanwhile itself.  After its last tile a wave keeps
// trying (up to 64 tiles a hold) until every pair has every tile.  Both
// waits end once the tiles in flight are stored; they are bounded by
// wall-clock time (kDrainTicks of the 100 MHz clock), past which the
// workgroup is marked broken and the launch reports VGX_ETIMEOUT.
constexpr int kGroupAxes = 2;  // axes per adder lock (the two halves of a record word; one axis per lock was slower)
constexpr int kPairs = (VGX_MAXDIM + kGroupAxes - 1) / kGroupAxes;
constexpr uint32_t kRing = 128;
struct HistSeq {
    uint32_t done[kRing];  // tile n complete (records stored): done[n % kRing] == n + 1
    uint32_t word[kPairs];  // per pair: next tile << 1 | 1 while a wave adds
};
constexpr uint32_t kHoldTiles = 8;  // tiles per hold of a lock while the holder has tiles left to compute
constexpr uint32_t kBatch = 4;      // tiles whose records a holder loads at once
constexpr uint32_t kMinRun = 4;     // complete tiles a pair needs before a wave computing tiles takes its lock
constexpr uint64_t kDrainTicks = 3000000000ull;  // 30 s
// s_sleep argument between polls of a waiting wave (64 clocks each): a poll's
// instructions and LDS reads take issue slots from the waves still computing
constexpr int kPollSleep = 16;

// One lane's records of one tile for one axis pair: its 4 consecutive
// points' bin words and (w f n)^2.
struct PairRec {
    uint4 wd;
    double2 c01, c23;
};

// Loads the lane's records of axis pair k for the tile at batch point ri0
// (16-byte loads: the rows are padded to whole tiles, so they stay inside
// them).
__device__ __forceinline__ void pair_load(const TileIO& io, int k, uint32_t ri0, int lane, PairRec& rc) {
    const uint32_t r0 = ri0 + (uint32_t)lane * kK;
    rc.wd = *reinterpret_cast<const uint4*>(io.rec_bins + (size_t)((k * kGroupAxes) >> 1) * io.rec_stride + r0);
    rc.c01 = *reinterpret_cast<const double2*>(io.rec_c2 + r0);
    rc.c23 = *reinterpret_cast<const double2*>(io.rec_c2 + r0 + 2);
}

// The loaded records of a tile with nvalid points added to the LDS sums of
// axes 2k and 2k+1 (< dim); records past the tile's last point are skipped.
__device__ __forceinline__ void pair_add(const TileGeom& g, int dim, int k, const PairRec& rc, uint32_t nvalid,
                                         double* Ls, int lane) {
    const uint32_t nmine = (uint32_t)lane * kK < nvalid ? min((uint32_t)kK, nvalid - (uint32_t)lane * kK) : 0u;
    const double cv[kK] = {rc.c01.x, rc.c01.y, rc.c23.x, rc.c23.y};
    const uint32_t ws[kK] = {rc.wd.x, rc.wd.y, rc.wd.z, rc.wd.w};
#pragma unroll
    for (int h = 0; h < kGroupAxes; ++h) {
        const int d = kGroupAxes * k + h;
        if (d >= dim) break;
#pragma unroll
        for (int j = 0; j < kK; ++j) {
            const uint32_t kb = (ws[j] >> (16 * (d & 1))) & 0xFFFFu;
            if ((uint32_t)j < nmine && kb != kNoBin) unsafeAtomicAdd(&Ls[d * g.nb + (int)kb], cv[j]);
        }
    }
}

// The number of consecutive complete tiles from tile n on, at most cap
// (<= 64): lane i looks at tile n + i.
__device__ __forceinline__ uint32_t ready_run(HistSeq& hs, uint32_t n, uint32_t ntile, uint32_t cap, int lane) {
    const uint32_t t = n + (uint32_t)lane;
    bool ok = false;
    if ((uint32_t)lane < cap && t < ntile)
        ok = __hip_atomic_load(&hs.done[t % kRing], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == t + 1u;
    const uint64_t run = ~__ballot(ok);
    return run ? (uint32_t)__builtin_ctzll(run) : 64u;
}

// One try of the calling wave on axis pair k: from the pair's next tile on,
// up to cap (<= 64) tiles in order while their records are complete.  The
// pair's word holds its next tile and the lock bit, so taking the lock also
// proves that no other wave has added those tiles since the snapshot.
// ntile tiles in this workgroup from batch tile tb; npts points in the
// batch.  Wave-uniform.
__device__ __forceinline__ void pair_try_add(const TileGeom& g, const TileIO& io, int dim, int k, HistSeq& hs,
                                             uint32_t tb, uint32_t ntile, uint32_t npts, double* Ls, int lane,
                                             uint32_t cap) {
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(&hs.word[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (w0 & 1u) return;  // another wave adds
    const uint32_t n = w0 >> 1;
    const uint32_t m = __builtin_amdgcn_readfirstlane(ready_run(hs, n, ntile, cap, lane));
    if (m == 0) return;
    // the records of complete tiles never change: the first batch is loaded
    // before the lock is taken
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    PairRec R[kBatch];
#pragma unroll
    for (uint32_t q = 0; q < kBatch; ++q)
        if (q < m) pair_load(io, k, (tb + n + q) * (uint32_t)kT, lane, R[q]);
    uint32_t got = 0u;
    if (lane == 0) {
        uint32_t expect = w0;
        got = __hip_atomic_compare_exchange_strong(&hs.word[k], &expect, w0 | 1u, __ATOMIC_ACQUIRE,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  ? 1u
                  : 0u;
    }
    got = __builtin_amdgcn_readfirstlane(got);
    if (!got) return;
#pragma unroll 1
    for (uint32_t i = 0;;) {
        const uint32_t nb = min(kBatch, m - i);
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q) {
            if (q < nb) {
                const uint32_t ri0 = (tb + n + i + q) * (uint32_t)kT;
                pair_add(g, dim, k, R[q], min((uint32_t)kT, npts - ri0), Ls, lane);
            }
        }
        i += nb;
        if (i >= m) break;
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            if (i + q < m) pair_load(io, k, (tb + n + i + q) * (uint32_t)kT, lane, R[q]);
    }
    // the additions are complete before the next holder's
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(&hs.word[k], (n + m) << 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One try on every axis pair that is free and has its next tile complete
// (and the one min_run - 1 tiles later: while tiles are still computed the
// locks are taken for runs of several tiles), from pair `first` on; true
// when every pair had every tile at the look.
__device__ __forceinline__ bool hist_try_add(const TileGeom& g, const TileIO& io, int dim, HistSeq& hs, uint32_t tb,
                                             uint32_t ntile, uint32_t npts, double* Ls, int lane, int first,
                                             uint32_t cap, uint32_t min_run) {
    const int np = (dim + kGroupAxes - 1) / kGroupAxes;
    // one look at every pair (lane k): finished, or free with its next tile
    // complete (and tile min_run - 1 after it, or the last tile); only the
    // latter are tried
    bool fin = true, open = false;
    if (lane < np) {
        const uint32_t w = __hip_atomic_load(&hs.word[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t nx = w >> 1;
        fin = nx >= ntile;
        const uint32_t nl = min(nx + min_run, ntile) - 1u;  // needs !fin
        open = !(w & 1u) && !fin &&
               __hip_atomic_load(&hs.done[nx % kRing], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == nx + 1u &&
               __hip_atomic_load(&hs.done[nl % kRing], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == nl + 1u;
    }
    if (__ballot(!fin) == 0) return true;
    const uint64_t want = __ballot(open);
#pragma unroll 1
    for (int i = 0; i < np; ++i) {
        int k = first + i;
        k = k >= np ? k - np : k;
        if ((want >> k) & 1ull) pair_try_add(g, io, dim, k, hs, tb, ntile, npts, Ls, lane, cap);
    }
    return false;
}

// Whether every pair has added every tile before tile `upto`.
__device__ __forceinline__ bool added_before(HistSeq& hs, int dim, uint32_t upto, int lane) {
    bool ok = true;
    if (lane < (dim + kGroupAxes - 1) / kGroupAxes)
        ok = (__hip_atomic_load(&hs.word[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> 1) >= upto;
    return __ballot(!ok) == 0;
}

// Workgroup reduction of the wave accumulators into out[0..5] (DD mean,
// var, sigf sums, fixed order) and the run-wide max |w f n| / NaN flag.
__device__ inline void block_close(TileAcc& acc, double* Lred, double* out, Scalars* sc) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
    acc.am = wave_sum_dd(acc.am);
    acc.av = wave_sum_dd(acc.av);
    acc.as = wave_sum_dd(acc.as);
    const double wmax = wave_max(acc.wmax);
    const int anynan = __any(acc.nan_seen);
    if (lane == 0) {
        Lred[wave * 6 + 0] = acc.am.hi; Lred[wave * 6 + 1] = acc.am.lo;
        Lred[wave * 6 + 2] = acc.av.hi; Lred[wave * 6 + 3] = acc.av.lo;
        Lred[wave * 6 + 4] = acc.as.hi; Lred[wave * 6 + 5] = acc.as.lo;
        if (wmax > 0.0) atomicMax((unsigned long long*)&sc->wfn_max_bits,
                                  (unsigned long long)__double_as_longlong(wmax));
        if (anynan) atomicOr(&sc->status, (int32_t)VGX_ENAN);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD m{0, 0}, v{0, 0}, s{0, 0};
        for (int k = 0; k < nwave; ++k) {
            m = dd_sum(m, DD{Lred[k * 6 + 0], Lred[k * 6 + 1]});
            v = dd_sum(v, DD{Lred[k * 6 + 2], Lred[k * 6 + 3]});
            s = dd_sum(s, DD{Lred[k * 6 + 4], Lred[k * 6 + 5]});
        }
        out[0] = m.hi; out[1] = m.lo; out[2] = v.hi; out[3] = v.lo; out[4] = s.hi; out[5] = s.lo;
    }
}

// Workgroup close of the fused sampler: the sums of its tiles
// tile_part[tb..te) in tile order, in DD, into out[0..5] (mean, var, sigf
// sums; lane l takes tiles tb + l, tb + l + 64, ..., then a fixed-order
// butterfly), and
// the run-wide max |w f n| / NaN flag from the lanes' accumulators.
__device__ inline void block_close_tiles(const TileAcc& acc, const double* tpart, uint32_t tb, uint32_t te,
                                         double* out, Scalars* sc) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double wmax = wave_max(acc.wmax);
    const int anynan = __any(acc.nan_seen);
    if (lane == 0) {
        if (wmax > 0.0) atomicMax((unsigned long long*)&sc->wfn_max_bits,
                                  (unsigned long long)__double_as_longlong(wmax));
        if (anynan) atomicOr(&sc->status, (int32_t)VGX_ENAN);
    }
    __syncthreads();  // every tile's sums are stored
    if (wave == 0) {
        DD m{0, 0}, v{0, 0}, q{0, 0};
        for (uint32_t t = tb + (uint32_t)lane; t < te; t += 64) {
            const double* tp = tpart + 3 * (size_t)t;
            dd_add(m, tp[0]);
            dd_add(v, tp[1]);
            dd_add(q, tp[2]);
        }
        m = wave_sum_dd(m);
        v = wave_sum_dd(v);
        q = wave_sum_dd(q);
        if (lane == 0) {
            out[0] = m.hi; out[1] = m.lo; out[2] = v.hi; out[3] = v.lo; out[4] = q.hi; out[5] = q.lo;
        }
    }
}

template <int DIM, bool PER, class Fcn>
__global__ __launch_bounds__(kFusedThreads) void fused_kernel(IterGeom g, FusedPtrs io, Fcn fcn) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_skip;
    __shared__ uint32_t s_broken;
    __shared__ uint32_t s_tile;  // the workgroup's next tile to take
    __shared__ HistSeq hs;
    Scalars* sc = io.sc;
    // NaN seen earlier in this run: nothing more is done (read once, so that
    // every wave of the workgroup takes the same decision)
    if (threadIdx.x == 0) {
        s_skip = sc->status;
        s_broken = 0u;
        s_tile = 0u;
    }
    if (threadIdx.x < kRing) hs.done[threadIdx.x] = 0u;
    if (threadIdx.x < kPairs) hs.word[threadIdx.x] = 0u;
    __syncthreads();
    if (s_skip) return;
    const TileGeom tg = tile_geom(g);
    const int dim = DIM > 0 ? DIM : g.dim;
    const uint32_t ntiles = sc->b_ntiles, p0 = sc->b_p0, p1 = sc->b_p1, hend = sc->b_h1;
    const int tot = dim * g.nb;
    // the ordered LDS histogram is compiled into the integer stratum path (the
    // fast FP32 sampler) only; the other paths train from their records
    // after the launch
    const bool lds_hist = PER && g.hist_lds && g.train_points;

    // ---- LDS carve-up: grid | histogram sums, u16 pair counts | offsets
    // slices
    unsigned char* cur = smem;
    const AxisC* ax = io.axes;  // read by scalar loads (ldc), not staged
    double* Lg = reinterpret_cast<double*>(cur);
    for (int i = threadIdx.x; i < dim * g.ld; i += blockDim.x) Lg[i] = io.grid[i];
    const double* G = Lg;
    cur += (size_t)dim * g.ld * sizeof(double);
    double* Ls = nullptr;
    uint32_t* Lc = nullptr;
    if (lds_hist) {
        Ls = reinterpret_cast<double*>(cur);
        cur += (size_t)tot * sizeof(double);
        Lc = reinterpret_cast<uint32_t*>(cur);
        cur += (size_t)((tot + 1) / 2) * sizeof(uint32_t);
        for (int i = threadIdx.x; i < tot; i += blockDim.x) Ls[i] = 0.0;
        for (int i = threadIdx.x; i < (tot + 1) / 2; i += blockDim.x) Lc[i] = 0u;
    }
    // offsets from smem rounded up to 16 B (pointer arithmetic on smem keeps
    // the accesses ds_* instead of flat)
    size_t o16 = ((size_t)(cur - smem) + 15) & ~(size_t)15;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nwave = blockDim.x >> 6;
    uint32_t* offs = reinterpret_cast<uint32_t*>(smem + o16) + wave * kNoff;

    // unordered histogram: the tile body adds each point into the LDS
    // histogram at once (no records, no fixed order)
    const bool unordered = lds_hist && g.hist_unordered;
    const TileIO tio{io.off, io.sigf, io.slots, io.errv, unordered ? nullptr : io.rec_bins,
                     unordered ? nullptr : io.rec_c2, p0, io.rec_stride};
    double* const Lsu = unordered ? Ls : nullptr;
    uint32_t* const Lcu = lds_hist ? Lc : nullptr;
    const uint32_t chunk = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint32_t tb = blockIdx.x * chunk;
    const uint32_t te = min(tb + chunk, ntiles);
    const uint32_t ntile_wg = te > tb ? te - tb : 0u;
    const uint32_t off_end = io.off[hend];
    // ordered: the tiles enter the histogram in tile order through the pair
    // locks (hist_try_add), from the records the tile body stores
    const bool ordered = lds_hist && !unordered;
    const int ngroups = (dim + kGroupAxes - 1) / kGroupAxes;  // adder locks (axis groups)
    uint32_t* const crow = lds_hist ? io.hpart_c + (size_t)blockIdx.x * tot : nullptr;
    // u16 LDS counts: a workgroup with more than kCntFlushTiles tiles moves
    // them into its u32 row (zeroed here) every kCntFlushTiles tiles taken,
    // before any can overflow: between two moves at most kCntFlushTiles + 15
    // tiles (15 more waves' tiles in flight) add to a bin, 256 points each
    const bool cnt_moves = lds_hist && ntile_wg > kCntFlushTiles;
    if (cnt_moves) {
        for (int i = threadIdx.x; i < tot; i += blockDim.x)
            __hip_atomic_store(&crow[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
    }
    __syncthreads();

    TileAcc acc{{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, 0.0, 0};
#pragma unroll 1
    for (;;) {
        uint32_t n = 0u;
        if (lane == 0) n = atomicAdd(&s_tile, 1u);
        n = __builtin_amdgcn_readfirstlane(n);
        if (n >= ntile_wg) break;
        if constexpr (PER) {
            if (ordered && n >= kRing) {
                // tile n - kRing's flag slot is reused: every pair has added it
                const uint64_t t_wait0 = wall_clock64();
#pragma unroll 1
                while (!added_before(hs, dim, n - kRing + 1u, lane)) {
                    if (__builtin_amdgcn_readfirstlane(
                            __hip_atomic_load(&s_broken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)))
                        break;
                    (void)hist_try_add(tg, tio, dim, hs, tb, ntile_wg, p1 - p0, Ls, lane, wave % ngroups,
                                       64u, 1u);
                    if (wall_clock64() - t_wait0 > kDrainTicks) {
                        if (lane == 0)
                            __hip_atomic_store(&s_broken, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(kPollSleep);
                }
                if (__builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(&s_broken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)))
                    break;
            }
            if (cnt_moves && n % kCntFlushTiles == kCntFlushTiles - 1) {
                // each word taken atomically: counts added meanwhile stay
#pragma unroll 1
                for (int i = lane; i < (tot + 1) / 2; i += 64) {
                    const uint32_t wv = atomicExch(&Lc[i], 0u);
                    if (wv & 0xffffu) atomicAdd(&crow[2 * i], wv & 0xffffu);
                    if ((wv >> 16) && 2 * i + 1 < tot) atomicAdd(&crow[2 * i + 1], wv >> 16);
                }
            }
        }
        const uint32_t t = tb + n;
        const uint32_t plo = p0 + t * kT;
        const TilePos tp{t, plo, min(plo + (uint32_t)kT, p1), (uint32_t)io.tile_h[t], hend, off_end, 0u, 0u};
        uint32_t ov[kOffRegs];
        load_offsets(io.off, tp.hf, hend, off_end, lane, ov);
        store_offsets(offs, lane, ov);
        tile_body<DIM, PER>(tg, tio, fcn, ax, G, Lsu, Lcu, offs, tp, lane, acc);
        {
            // the tile's sums (in lane 63): a fixed FP64 tree over the lanes'
            // values (each at most kK hcubes' terms); summed in DD in tile
            // order at the close
            const double m = wave_total(acc.am.hi + acc.am.lo), v = wave_total(acc.av.hi + acc.av.lo),
                         q = wave_total(acc.as.hi + acc.as.lo);
            if (lane == 63) {
                double* o = io.tile_part + 3 * (size_t)t;
                o[0] = m; o[1] = v; o[2] = q;
            }
            acc.am = DD{0.0, 0.0};
            acc.av = DD{0.0, 0.0};
            acc.as = DD{0.0, 0.0};
        }
        if constexpr (PER) {
            if (ordered) {
                // the tile's records are stored before it is marked complete
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lane == 0)
                    __hip_atomic_store(&hs.done[n % kRing], n + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                (void)hist_try_add(tg, tio, dim, hs, tb, ntile_wg, p1 - p0, Ls, lane, wave % ngroups,
                                   kHoldTiles, kMinRun);
            }
        }
    }
    if constexpr (PER) {
        if (ordered) {
            // every wave helps until the workgroup's last tile is in
            const uint64_t t_wait0 = wall_clock64();
#pragma unroll 1
            for (;;) {
                if (__builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(&s_broken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)))
                    break;
                if (hist_try_add(tg, tio, dim, hs, tb, ntile_wg, p1 - p0, Ls, lane, wave % ngroups, 64u, 1u))
                    break;
                if (wall_clock64() - t_wait0 > kDrainTicks) {
                    if (lane == 0) __hip_atomic_store(&s_broken, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    break;
                }
                __builtin_amdgcn_s_sleep(kPollSleep);
            }
        }
    }

    // ---- block epilogue ----
    __syncthreads();
    if (lds_hist) {
        // this workgroup's row of the histogram: plain stores (counts added
        // to the row when they were moved there before), summed in
        // workgroup order after the launch
        double* srow = io.hpart_s + (size_t)blockIdx.x * tot;
        for (int i = threadIdx.x; i < tot; i += blockDim.x) srow[i] = Ls[i];
        if (cnt_moves) {
            for (int i = threadIdx.x; i < (tot + 1) / 2; i += blockDim.x) {
                const uint32_t wv = Lc[i];
                if (wv & 0xffffu) atomicAdd(&crow[2 * i], wv & 0xffffu);
                if ((wv >> 16) && 2 * i + 1 < tot) atomicAdd(&crow[2 * i + 1], wv >> 16);
            }
        } else {
            counts_to_row(Lc, tot, crow, false, false);
        }
        if (threadIdx.x == 0 && s_broken) atomicCAS(&sc->status, 0, (int32_t)VGX_ETIMEOUT);
    }
    block_close_tiles(acc, io.tile_part, tb, te, io.blk_part + 6 * (size_t)blockIdx.x, sc);
}

template <int DIM, class Fcn>
int launch_fused_dim(const IterGeom& g, const FusedPtrs& io, const Fcn& fcn, int blocks, size_t lds,
                     hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    auto kern = fused_kernel<DIM, false, Fcn>;
    if constexpr (Fcn::kRng32) {
        if (g.per_all) kern = fused_kernel<DIM, true, Fcn>;
    }
    // the LDS limit of this instantiation is raised once, to the largest
    // request seen (a host-side attribute, not a per-launch cost)
    static size_t lds_set = 64 * 1024;
    if (lds > lds_set) {
        if (hipFuncSetAttribute((const void*)fused_kernel<DIM, false, Fcn>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return VGX_EHIP;
        if constexpr (Fcn::kRng32) {
            if (hipFuncSetAttribute((const void*)fused_kernel<DIM, true, Fcn>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
                return VGX_EHIP;
        }
        lds_set = lds;
    }
    if (ev0)
        hipExtLaunchKernelGGL(kern, dim3(blocks), dim3(kFusedThreads), (uint32_t)lds, st, ev0, ev1, 0, g, io, fcn);
    else
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kFusedThreads), lds, st, g, io, fcn);
    return check_launch();
}

}  // namespace vgx
