// vgx_hist.hip -- AdaptiveMap.add_training_data (src/vegas/_vegas.pyx:421-464).
//
// Fast mode: each workgroup privatises the (dim x ninc) histogram in LDS
// (FP64 sums via ds_add_f64, u32 counts), then flushes it with one global
// FP64 atomic per touched bin.  Counts are exact; sums equal the
// reference's up to the order of additions.
//
// Ordered mode: bit-for-bit the reference's sequential loop.  A stable
// counting sort groups |f| by (axis, bin) keeping point order, then one lane
// per bin adds the group in point order onto the existing sum:
//   1 hist_count    per-tile, per-bin counts (LDS)
//   2 hist_scan     exclusive scan over tiles per bin, then over bins
//   3 hist_scatter  one wave per (tile, axis); stable ranks from ballots
//   4 hist_sum      sequential per-bin sums in point order
#include <mutex>

#include "vgx_internal.h"

namespace vgx {

constexpr int kHistTile = 1024;  // points per tile (ordered mode)

struct HistArgs {
    int dim;
    int nb;  // ld - 1: row stride of sum_f / n_f
    int use_lds;
    int ntiles;
    int32_t ninc[VGX_MAXDIM];
    const double* y;
    const double* f;
    int64_t n;
    double* sum_f;
    double* n_f;
    uint32_t* cnt;     // ordered: dim x ntiles x nb
    uint32_t* start;   // ordered: dim x nb bin starts
    uint32_t* total;   // ordered: dim x nb bin totals
    double* sorted;    // ordered: dim x n
    // training records of the fused sampler instead of y and f: the bin of
    // point i on axis d in the (d & 1) half of rbins[(d >> 1) * rstride + i],
    // the values in f, the point count read on the device as nrange[1] -
    // nrange[0] (n is then the capacity)
    const uint32_t* rbins;
    int64_t rstride;
    const uint32_t* nrange;
};

__device__ inline int64_t npoints(const HistArgs& a) {
    return a.nrange ? (int64_t)(a.nrange[1] - a.nrange[0]) : a.n;
}

// Bin of y for axis d, or -1 when y is not strictly inside (0, 1).
__device__ inline int train_bin(double yv, int ni) {
    if (!(yv > 0 && yv < 1)) return -1;
    int k = (int)floor(yv * (double)ni);
    // y*ninc can round up to ninc for y within an ulp of 1; the reference
    // would then write one slot past the row.  Keep it in the last bin.
    return k < ni ? k : ni - 1;
}

// Bin of point i on axis d: from its record, or from y (pyx:459-463).
__device__ inline int point_bin(const HistArgs& a, int64_t i, int d, int ni) {
    if (a.rbins) {
        const uint32_t b = (a.rbins[(d >> 1) * a.rstride + i] >> (16 * (d & 1))) & 0xFFFFu;
        return (b == kNoBin || (int)b >= ni) ? -1 : (int)b;  // (a bin >= ninc is never written)
    }
    return train_bin(a.y[i * a.dim + d], ni);
}

__global__ __launch_bounds__(256) void hist_fast_kernel(HistArgs a) {
    extern __shared__ double lds[];
    double* s = lds;
    uint32_t* c = reinterpret_cast<uint32_t*>(lds + (a.use_lds ? a.dim * a.nb : 0));
    const int tot = a.dim * a.nb;
    if (a.use_lds) {
        for (int i = threadIdx.x; i < tot; i += blockDim.x) {
            s[i] = 0.0;
            c[i] = 0u;
        }
        __syncthreads();
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double fa = fabs(a.f[i]);
        for (int d = 0; d < a.dim; ++d) {
            int k = train_bin(a.y[i * a.dim + d], a.ninc[d]);
            if (k < 0) continue;
            if (a.use_lds) {
                unsafeAtomicAdd(&s[d * a.nb + k], fa);
                atomicAdd(&c[d * a.nb + k], 1u);
            } else {
                unsafeAtomicAdd(&a.sum_f[d * a.nb + k], fa);
                unsafeAtomicAdd(&a.n_f[d * a.nb + k], 1.0);
            }
        }
    }
    if (!a.use_lds) return;
    __syncthreads();
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        if (c[i]) {
            unsafeAtomicAdd(&a.sum_f[i], s[i]);
            unsafeAtomicAdd(&a.n_f[i], (double)c[i]);
        }
    }
}

// ---- ordered mode ---------------------------------------------------------
__global__ __launch_bounds__(256) void hist_count_kernel(HistArgs a) {
    extern __shared__ uint32_t cl[];
    const int tile = blockIdx.x, d = blockIdx.y;
    const int ni = a.ninc[d];
    for (int k = threadIdx.x; k < a.nb; k += blockDim.x) cl[k] = 0u;
    __syncthreads();
    const int64_t p0 = (int64_t)tile * kHistTile, n = npoints(a);
    for (int j = threadIdx.x; j < kHistTile; j += blockDim.x) {
        int64_t i = p0 + j;
        if (i >= n) break;
        int k = point_bin(a, i, d, ni);
        if (k >= 0) atomicAdd(&cl[k], 1u);
    }
    __syncthreads();
    uint32_t* out = a.cnt + ((int64_t)d * a.ntiles + tile) * a.nb;
    for (int k = threadIdx.x; k < a.nb; k += blockDim.x) out[k] = cl[k];
}

// One workgroup per axis: per-bin exclusive scan over tiles (in place), then
// an exclusive scan over bins of the bin totals.
__global__ __launch_bounds__(1024) void hist_scan_kernel(HistArgs a) {
    __shared__ uint32_t part[1024];
    const int d = blockIdx.x;
    for (int k = threadIdx.x; k < a.nb; k += blockDim.x) {
        uint32_t run = 0;
        for (int t = 0; t < a.ntiles; ++t) {
            uint32_t* p = a.cnt + ((int64_t)d * a.ntiles + t) * a.nb + k;
            uint32_t v = *p;
            *p = run;
            run += v;
        }
        a.total[d * a.nb + k] = run;
    }
    __syncthreads();
    // exclusive scan of total[d][0..nb) -> start[d][..]; chunked over 1024 bins
    uint32_t carry = 0;
    for (int base = 0; base < a.nb; base += blockDim.x) {
        int k = base + threadIdx.x;
        uint32_t v = k < a.nb ? a.total[d * a.nb + k] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            uint32_t add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (k < a.nb) a.start[d * a.nb + k] = carry + part[threadIdx.x] - v;
        uint32_t chunk_total = part[blockDim.x - 1];
        __syncthreads();
        carry += chunk_total;
    }
}

// One wave per (tile, axis).  Points of the tile are taken 64 at a time in
// order; equal keys within a group get consecutive ranks in lane order.
__global__ __launch_bounds__(64) void hist_scatter_kernel(HistArgs a) {
    extern __shared__ uint32_t run[];
    const int tile = blockIdx.x, d = blockIdx.y;
    const int ni = a.ninc[d];
    const int lane = threadIdx.x;
    for (int k = lane; k < a.nb; k += 64) run[k] = 0u;
    __syncthreads();
    const uint32_t* tbase = a.cnt + ((int64_t)d * a.ntiles + tile) * a.nb;
    const uint32_t* bstart = a.start + d * a.nb;
    double* out = a.sorted + (int64_t)d * a.n;
    const int64_t p0 = (int64_t)tile * kHistTile, n = npoints(a);
    for (int g = 0; g < kHistTile; g += 64) {
        const int64_t i = p0 + g + lane;
        int key = -1;
        double fa = 0.0;
        if (i < n) {
            key = point_bin(a, i, d, ni);
            fa = fabs(a.f[i]);
        }
        uint64_t active = __ballot(key >= 0);
        while (active) {
            const int leader = __ffsll((unsigned long long)active) - 1;
            const int lk = __shfl(key, leader, 64);
            const uint64_t m = __ballot(key == lk) & active;
            const uint32_t base = run[lk];
            if ((m >> lane) & 1ull) {
                const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                out[bstart[lk] + tbase[lk] + base + rank] = fa;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == leader) run[lk] = base + (uint32_t)__popcll(m);
            __builtin_amdgcn_wave_barrier();
            active &= ~m;
        }
    }
}

__global__ __launch_bounds__(256) void hist_sum_kernel(HistArgs a) {
    const int d = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.ninc[d]) return;
    const uint32_t s = a.start[d * a.nb + k], t = a.total[d * a.nb + k];
    const double* in = a.sorted + (int64_t)d * a.n;
    double acc = a.sum_f[d * a.nb + k];
    for (uint32_t j = 0; j < t; ++j) acc += in[s + j];
    a.sum_f[d * a.nb + k] = acc;
    a.n_f[d * a.nb + k] = a.n_f[d * a.nb + k] + (double)t;
}

}  // namespace vgx

using namespace vgx;

extern "C" int vgx_train_hist(vgx_ctx* ctx, int dim, const int32_t* ninc, int ld, const double* y,
                              const double* f, int64_t n, double* sum_f, double* n_f, int ordered,
                              void* stream) {
    if (dim < 1 || dim > VGX_MAXDIM || ld < 2 || n < 0 || !ninc) return VGX_EINVAL;
    if (n == 0) return VGX_OK;
    if (n > 0xFFFFFFFFll) return VGX_ERANGE;
    hipStream_t st = (hipStream_t)stream;
    HistArgs a{};
    a.dim = dim;
    a.nb = ld - 1;
    a.y = y;
    a.f = f;
    a.n = n;
    a.sum_f = sum_f;
    a.n_f = n_f;
    for (int d = 0; d < dim; ++d) {
        if (ninc[d] < 1 || ninc[d] > a.nb) return VGX_EINVAL;
        a.ninc[d] = ninc[d];
    }
    if (!ordered) {
        size_t bytes = (size_t)dim * a.nb * (sizeof(double) + sizeof(uint32_t));
        a.use_lds = bytes <= 64 * 1024;
        int64_t b = (n + 255) / 256;
        if (b > 512) b = 512;
        hipLaunchKernelGGL(hist_fast_kernel, dim3((unsigned)b), dim3(256), a.use_lds ? bytes : 0,
                           st, a);
        return check_launch();
    }
    if (!ctx) return VGX_EINVAL;
    if ((size_t)a.nb * sizeof(uint32_t) > 64 * 1024) return VGX_ERANGE;
    a.ntiles = (int)((n + kHistTile - 1) / kHistTile);
    size_t ncnt = (size_t)dim * a.ntiles * a.nb;
    int rc = ctx->hist_cnt.ensure(ncnt * sizeof(uint32_t));
    if (rc) return rc;
    rc = ctx->hist_start.ensure(2 * (size_t)dim * a.nb * sizeof(uint32_t));
    if (rc) return rc;
    rc = ctx->hist_sorted.ensure((size_t)dim * n * sizeof(double));
    if (rc) return rc;
    a.cnt = (uint32_t*)ctx->hist_cnt.ptr;
    a.start = (uint32_t*)ctx->hist_start.ptr;
    a.total = a.start + (size_t)dim * a.nb;
    a.sorted = (double*)ctx->hist_sorted.ptr;
    size_t lds = (size_t)a.nb * sizeof(uint32_t);
    hipLaunchKernelGGL(hist_count_kernel, dim3(a.ntiles, dim), dim3(256), lds, st, a);
    hipLaunchKernelGGL(hist_scan_kernel, dim3(dim), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(hist_scatter_kernel, dim3(a.ntiles, dim), dim3(64), lds, st, a);
    int maxn = 0;
    for (int d = 0; d < dim; ++d) maxn = ninc[d] > maxn ? ninc[d] : maxn;
    hipLaunchKernelGGL(hist_sum_kernel, dim3((maxn + 255) / 256, dim), dim3(256), 0, st, a);
    return check_launch();
}

namespace vgx {

// The training histogram of a batch of the fused sampler, from its records,
// bit for bit the reference's add_training_data on the batch's points in
// order (pyx:457-463): the ordered mode above with the bins and values of
// the records.  cap bounds the batch's point count, which is read on the
// device from nrange.
int train_hist_records(vgx_ctx* ctx, int dim, const int32_t* ninc, int nb, const uint32_t* rbins,
                       int64_t rstride, const double* c2, const uint32_t* nrange, int64_t cap, double* sum_f,
                       double* n_f, hipStream_t st) {
    if (cap <= 0) return VGX_OK;
    if ((size_t)nb * sizeof(uint32_t) > 64 * 1024 || dim > VGX_MAXDIM) return VGX_ERANGE;
    HistArgs a{};
    a.dim = dim;
    a.nb = nb;
    a.n = cap;
    a.f = c2;
    a.sum_f = sum_f;
    a.n_f = n_f;
    a.rbins = rbins;
    a.rstride = rstride;
    a.nrange = nrange;
    int maxn = 0;
    for (int d = 0; d < dim; ++d) {
        a.ninc[d] = ninc[d];
        maxn = ninc[d] > maxn ? ninc[d] : maxn;
    }
    a.ntiles = (int)((cap + kHistTile - 1) / kHistTile);
    int rc = ctx->hist_cnt.ensure((size_t)dim * a.ntiles * a.nb * sizeof(uint32_t));
    if (rc) return rc;
    if ((rc = ctx->hist_start.ensure(2 * (size_t)dim * a.nb * sizeof(uint32_t)))) return rc;
    if ((rc = ctx->hist_sorted.ensure((size_t)dim * cap * sizeof(double)))) return rc;
    a.cnt = (uint32_t*)ctx->hist_cnt.ptr;
    a.start = (uint32_t*)ctx->hist_start.ptr;
    a.total = a.start + (size_t)dim * a.nb;
    a.sorted = (double*)ctx->hist_sorted.ptr;
    const size_t lds = (size_t)a.nb * sizeof(uint32_t);
    hipLaunchKernelGGL(hist_count_kernel, dim3(a.ntiles, dim), dim3(256), lds, st, a);
    hipLaunchKernelGGL(hist_scan_kernel, dim3(dim), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(hist_scatter_kernel, dim3(a.ntiles, dim), dim3(64), lds, st, a);
    hipLaunchKernelGGL(hist_sum_kernel, dim3((maxn + 255) / 256, dim), dim3(256), 0, st, a);
    return check_launch();
}

}  // namespace vgx

// ---- the ordered histogram without allocation -----------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): the ordered mode of vgx_train_hist with its
// temporaries (per-tile counts, bin starts and totals, the grouped values)
// in this file-scope array, one copy per GPU, held until the process ends.
// Its device address is looked up once per device (kernels take that
// address, never the array's host-side name).
namespace vgx {
constexpr size_t kNoAllocScratch = (size_t)128 << 20;  // bytes per device
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

constexpr int kNoAllocDevices = 64;
static void* g_scratch_dev[kNoAllocDevices] = {};
static std::mutex g_scratch_mu;

// The current device's copy of `scratch` (its device address).
static int scratch_for_current_device(unsigned char** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return VGX_EHIP;
    if (dev < 0 || dev >= kNoAllocDevices) return VGX_ERANGE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return VGX_EHIP;
        g_scratch_dev[dev] = p;
    }
    *out = (unsigned char*)g_scratch_dev[dev];
    return VGX_OK;
}

// Bytes of scratch the ordered histogram of n points needs, laid out as
// per-tile counts | bin starts and totals | grouped values (256-B aligned).
static size_t noalloc_layout(int dim, int nb, int64_t n, size_t* o_start, size_t* o_sorted) {
    const int64_t ntiles = (n + kHistTile - 1) / kHistTile;
    const size_t cnt = (((size_t)dim * ntiles * nb * sizeof(uint32_t)) + 255) & ~(size_t)255;
    const size_t st = ((2 * (size_t)dim * nb * sizeof(uint32_t)) + 255) & ~(size_t)255;
    *o_start = cnt;
    *o_sorted = cnt + st;
    return cnt + st + (size_t)dim * n * sizeof(double);
}
}  // namespace vgx

extern "C" int vgx_train_hist_noalloc(int dim, const int32_t* ninc, int ld, const double* y, const double* f,
                                      int64_t n, double* sum_f, double* n_f, void* stream) {
    if (dim < 1 || dim > VGX_MAXDIM || ld < 2 || n < 0 || !ninc) return VGX_EINVAL;
    if (n == 0) return VGX_OK;
    if (n > 0xFFFFFFFFll) return VGX_ERANGE;
    HistArgs a{};
    a.dim = dim;
    a.nb = ld - 1;
    a.y = y;
    a.f = f;
    a.n = n;
    a.sum_f = sum_f;
    a.n_f = n_f;
    int maxn = 0;
    for (int d = 0; d < dim; ++d) {
        if (ninc[d] < 1 || ninc[d] > a.nb) return VGX_EINVAL;
        a.ninc[d] = ninc[d];
        maxn = ninc[d] > maxn ? ninc[d] : maxn;
    }
    if ((size_t)a.nb * sizeof(uint32_t) > 64 * 1024) return VGX_ERANGE;
    size_t o_start = 0, o_sorted = 0;
    if (noalloc_layout(dim, a.nb, n, &o_start, &o_sorted) > kNoAllocScratch) return VGX_ERANGE;
    unsigned char* base = nullptr;
    int rc = scratch_for_current_device(&base);
    if (rc) return rc;
    a.ntiles = (int)((n + kHistTile - 1) / kHistTile);
    a.cnt = (uint32_t*)base;
    a.start = (uint32_t*)(base + o_start);
    a.total = a.start + (size_t)dim * a.nb;
    a.sorted = (double*)(base + o_sorted);
    hipStream_t st = (hipStream_t)stream;
    const size_t lds = (size_t)a.nb * sizeof(uint32_t);
    hipLaunchKernelGGL(hist_count_kernel, dim3(a.ntiles, dim), dim3(256), lds, st, a);
    hipLaunchKernelGGL(hist_scan_kernel, dim3(dim), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(hist_scatter_kernel, dim3(a.ntiles, dim), dim3(64), lds, st, a);
    hipLaunchKernelGGL(hist_sum_kernel, dim3((maxn + 255) / 256, dim), dim3(256), 0, st, a);
    return check_launch();
}
