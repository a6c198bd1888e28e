// vgx_ctx.hip -- context, error strings and device queries of libvgx.
#include <string.h>

#include "vgx_internal.h"

namespace vgx {

int device_info(DeviceInfo* out) {
    int dev = 0;
    VGX_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    VGX_HIP(hipGetDeviceProperties(&p, dev));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0) return VGX_ENODEV;
    out->device = dev;
    out->cus = p.multiProcessorCount;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        lds = 65536;
    out->max_lds = (size_t)lds;
    return VGX_OK;
}

}  // namespace vgx

extern "C" const char* vgx_strerror(int status) {
    switch (status) {
        case VGX_OK: return "ok";
        case VGX_EINVAL: return "invalid argument";
        case VGX_EHIP: return "HIP runtime error";
        case VGX_ENAN: return "integrand evaluates to nan";
        case VGX_ENOMEM: return "device allocation failed";
        case VGX_ERANGE: return "size beyond kernel limits";
        case VGX_ENODEV: return "no gfx950 device";
        case VGX_ETIMEOUT: return "a device wait timed out";
        default: return "unknown vgx status";
    }
}

extern "C" int vgx_version(void) { return 1; }

extern "C" int vgx_ctx_create(vgx_ctx** out) {
    if (!out) return VGX_EINVAL;
    vgx_ctx* c = new vgx_ctx();
    int rc = vgx::device_info(&c->info);
    if (rc) {
        delete c;
        return rc;
    }
    c->device = c->info.device;
    // timing-only events: no system-scope fence (cache writeback and
    // invalidate) when they complete, which would stall the next launch
    for (int i = 0; i < 2 * vgx_ctx::kMaxTimed; ++i) {
        if (hipEventCreateWithFlags(&c->ev[i], hipEventDisableSystemFence) != hipSuccess) {
            delete c;
            return VGX_EHIP;
        }
    }
    *out = c;
    return VGX_OK;
}

extern "C" int vgx_ctx_destroy(vgx_ctx* c) {
    if (!c) return VGX_OK;
    (void)hipDeviceSynchronize();
    vgx::Scratch* all[] = {&c->hist_cnt, &c->hist_sorted, &c->hist_start, &c->nh, &c->off,
                           &c->tile_h, &c->blocksum, &c->tile_slots, &c->tile_part, &c->blk_part,
                           &c->strad_part, &c->misc, &c->adapt_tmp, &c->errv, &c->cov_part,
                           &c->hist_part, &c->trec};
    if (vgx::g_timing_ctx == c) vgx::g_timing_ctx = nullptr;
    for (auto* s : all) s->release();
    for (int i = 0; i < 2 * vgx_ctx::kMaxTimed; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    delete c;
    return VGX_OK;
}

namespace vgx {
vgx_ctx* g_timing_ctx = nullptr;
}

extern "C" int vgx_set_timing(vgx_ctx* c, int on) {
    if (!c) return VGX_EINVAL;
    c->timing_on = on ? 1 : 0;
    c->ntimed = 0;
    if (on)
        vgx::g_timing_ctx = c;
    else if (vgx::g_timing_ctx == c)
        vgx::g_timing_ctx = nullptr;
    return VGX_OK;
}

extern "C" int vgx_timed_launches(vgx_ctx* c, float* ms, int32_t* tags, int cap) {
    if (!c || ((!ms || !tags) && cap > 0)) return VGX_EINVAL;
    const int n = c->ntimed < cap ? c->ntimed : cap;
    for (int i = 0; i < n; ++i) {
        VGX_HIP(hipEventSynchronize(c->ev[2 * i + 1]));
        VGX_HIP(hipEventElapsedTime(&ms[i], c->ev[2 * i], c->ev[2 * i + 1]));
        tags[i] = c->tags[i];
    }
    return n;
}

extern "C" int vgx_fused_times(vgx_ctx* c, float* ms, int cap) {
    if (!c || (!ms && cap > 0)) return VGX_EINVAL;
    int n = 0;
    for (int i = 0; i < c->ntimed && n < cap; ++i) {
        if (c->tags[i] != vgx::kTimedFused) continue;
        VGX_HIP(hipEventSynchronize(c->ev[2 * i + 1]));
        VGX_HIP(hipEventElapsedTime(&ms[n++], c->ev[2 * i], c->ev[2 * i + 1]));
    }
    return n;
}
