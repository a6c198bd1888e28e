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
    vgx::Scratch nh, off, tile_h, blocksum, tile_slots, blk_part, strad_part, misc, errv;
    // multi-component integrands: per-workgroup partials
    vgx::Scratch cov_part;
    vgx::Scratch hist_part;  // per-workgroup training-histogram rows of the fused sampler
    // adapt
    vgx::Scratch adapt_tmp;
    static constexpr int kMaxTimed = 256;
    hipEvent_t ev[2 * kMaxTimed] = {};
    int timing_on = 0;
    int ntimed = 0;
};
