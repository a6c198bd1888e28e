// vgx_internal.h -- host-side plumbing shared by the libvgx translation units.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "vgx_device.h"

namespace vgx {

inline int check_launch() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VGX_OK : VGX_EHIP;
}

#define VGX_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return VGX_EHIP;         \
    } while (0)

struct DeviceInfo {
    int device;
    int cus;
    size_t max_lds;  // bytes of LDS a workgroup may allocate
};

int device_info(DeviceInfo* out);

}  // namespace vgx
struct vgx_ctx;
namespace vgx {
// vgx_hist.hip: the training histogram of a fused-sampler batch from its
// records, in point order (bit for bit the reference's sequential sums).
int train_hist_records(vgx_ctx* ctx, int dim, const int32_t* ninc, int nb, const uint32_t* rbins,
                       int64_t rstride, const double* c2, const uint32_t* nrange, int64_t cap, double* sum_f,
                       double* n_f, hipStream_t st);

// A growable device buffer owned by a context.
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes) return VGX_OK;
        if (ptr) {
            if (hipFree(ptr) != hipSuccess) return VGX_EHIP;
            ptr = nullptr;
            bytes = 0;
        }
        size_t b = need < 256 ? 256 : need;
        b = (b + 255) & ~(size_t)255;
        if (hipMalloc(&ptr, b) != hipSuccess) return VGX_ENOMEM;
        bytes = b;
        return VGX_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

}  // namespace vgx

struct vgx_ctx {
    int device = 0;
    vgx::DeviceInfo info{};
    // ordered training histogram
    vgx::Scratch hist_cnt, hist_sorted, hist_start;
    // integrator
    vgx::Scratch nh, off, tile_h, blocksum, tile_slots, tile_part, blk_part, strad_part, misc, errv;
    // multi-component integrands: per-workgroup partials
    vgx::Scratch cov_part;
    vgx::Scratch hist_part;  // per-block training-histogram rows of the fused sampler
    vgx::Scratch trec;       // training records of the fused sampler (bins, (w f n)^2)
    // adapt
    vgx::Scratch adapt_tmp;
    static constexpr int kMaxTimed = 256;
    hipEvent_t ev[2 * kMaxTimed] = {};
    int32_t tags[kMaxTimed] = {};  // which kernel each event pair brackets (vgx::Timed*)
    int timing_on = 0;
    int ntimed = 0;
};

namespace vgx {
// Kernels whose launches can be timed with HIP events (vgx_set_timing).
enum Timed : int32_t {
    kTimedFused = 1,      // fused_kernel (single integral)
    kTimedBatchFused = 2, // b_fused_kernel (batched integrals, eager calls)
    kTimedFlowBwd = 3,    // flow_backward_kernel
    kTimedFlowWgrad = 4,  // (unused: the weight gradients are part of flow_backward_kernel)
    kTimedTrainRows = 5,  // train_rows_kernel
    kTimedFlowSample = 6, // flow_sample_kernel
    kTimedCov = 7         // the multi-component covariance pass
};
// The context whose timing is on (vgx_set_timing), for code without one.
extern vgx_ctx* g_timing_ctx;
// Events for the next launch of kernel `tag` when timing is on and the log
// has room (then true); the launch goes through hipExtLaunchKernelGGL.
inline bool timed_slot(vgx_ctx* c, int32_t tag, hipEvent_t* e0, hipEvent_t* e1) {
    if (!c || !c->timing_on || c->ntimed >= vgx_ctx::kMaxTimed) return false;
    *e0 = c->ev[2 * c->ntimed];
    *e1 = c->ev[2 * c->ntimed + 1];
    c->tags[c->ntimed] = tag;
    c->ntimed++;
    return true;
}
}  // namespace vgx
