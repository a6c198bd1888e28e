"""ctypes binding of libvgx.so (include/vgx.h).

The library is built in-tree (``vegas_amd/libvgx.so``) by ``vegas_amd.build``.
Nothing here falls back to CPU arithmetic: if the library or a gfx950 device
is missing, :func:`lib` raises and every kernel-backed call fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvgx.so")
MAXDIM = 64

VGX_OK, VGX_EINVAL, VGX_EHIP, VGX_ENAN, VGX_ENOMEM, VGX_ERANGE, VGX_ENODEV = 0, -1, -2, -3, -4, -5, -6
VGX_ETIMEOUT = -7
PREC_F64, PREC_F32 = 0, 1
FCN_GAUSSMIX, FCN_DIAGRIDGE, FCN_BUFFER = 1, 2, 3
FLAG_EXACT_ADAPT, FLAG_EXACT_DIV, FLAG_UNORDERED_HIST = 1, 2, 4


class VgxConfig(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ld", C.c_int32), ("ninc", C.c_int32 * MAXDIM),
                ("nstrat", C.c_int32 * MAXDIM), ("nhcube", C.c_int64), ("neval", C.c_int64),
                ("neval_frac", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("adapt", C.c_int32), ("adapt_to_errors", C.c_int32),
                ("min_neval_hcube", C.c_int64), ("max_neval_hcube", C.c_int64),
                ("seed", C.c_uint64), ("iteration", C.c_uint32), ("nitn", C.c_int32),
                ("precision", C.c_int32), ("flags", C.c_int32), ("weighted", C.c_int32),
                ("reserved", C.c_int32), ("uniforms", C.c_void_p)]


class VgxIntegrand(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ncomp", C.c_int32), ("A", C.c_void_p), ("B", C.c_void_p),
                ("c", C.c_void_p), ("fbuf", C.c_void_p), ("fbase", C.c_int64)]


class VgxState(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("sigf", C.c_void_p), ("sum_sigf", C.c_void_p),
                ("sum_f", C.c_void_p), ("n_f", C.c_void_p), ("itn_mean", C.c_void_p),
                ("itn_var", C.c_void_p), ("ravg", C.c_void_p), ("itn_neval", C.c_void_p),
                ("itn_nhrange", C.c_void_p), ("status", C.c_void_p)]


_lib = None
_ctx = {}


class VgxError(RuntimeError):
    pass


def lib():
    """Load libvgx.so; raise if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VgxError("libvgx.so not built: run `python -c 'import vegas_amd.build as b; b.build()'`")
    L = C.CDLL(LIB_PATH)
    vp, i32p = C.c_void_p, C.POINTER(C.c_int32)
    L.vgx_strerror.restype = C.c_char_p
    L.vgx_strerror.argtypes = [C.c_int]
    L.vgx_version.restype = C.c_int
    for nm in ("vgx_map_f64", "vgx_map_f32"):
        getattr(L, nm).argtypes = [C.c_int, i32p, vp, C.c_int, vp, vp, vp, C.c_int64, vp]
    L.vgx_invmap_f64.argtypes = [C.c_int, i32p, vp, C.c_int, vp, vp, vp, C.c_int64, vp]
    L.vgx_jac1d_f64.argtypes = [C.c_int, i32p, vp, C.c_int, vp, vp, C.c_int64, vp]
    L.vgx_ctx_create.argtypes = [C.POINTER(vp)]
    L.vgx_ctx_destroy.argtypes = [vp]
    L.vgx_train_hist.argtypes = [vp, C.c_int, i32p, C.c_int, vp, vp, C.c_int64, vp, vp, C.c_int, vp]
    L.vgx_adapt.argtypes = [vp, C.c_int, i32p, vp, C.c_int, C.c_int, vp, vp, C.c_double, i32p, vp,
                            C.c_int, C.c_int, vp]
    L.vgx_integrate.argtypes = [vp, C.POINTER(VgxConfig), C.POINTER(VgxIntegrand),
                                C.POINTER(VgxState), vp]
    L.vgx_iter_begin.argtypes = [vp, C.POINTER(VgxConfig), C.POINTER(VgxState), C.c_int, vp,
                                 C.POINTER(C.c_int64), vp]
    L.vgx_iter_sample.argtypes = [vp, C.POINTER(VgxConfig), C.POINTER(VgxState), C.c_int,
                                  C.c_int64, C.c_int64, C.c_int64, C.c_int64, vp, vp, vp, vp, vp]
    L.vgx_iter_accumulate.argtypes = [vp, C.POINTER(VgxConfig), C.POINTER(VgxState), C.c_int,
                                      C.c_int64, C.c_int64, C.c_int64, C.c_int64, vp, vp]
    L.vgx_iter_end.argtypes = [vp, C.POINTER(VgxConfig), C.POINTER(VgxState), C.c_int, vp]
    L.vgx_iter_accumulate_cov.argtypes = [vp, C.POINTER(VgxConfig), C.c_int64, C.c_int64, C.c_int64,
                                          C.c_int64, vp, vp, C.c_int, C.c_int, vp, vp, vp]
    L.vgx_set_timing.argtypes = [vp, C.c_int]
    L.vgx_fused_times.argtypes = [vp, C.POINTER(C.c_float), C.c_int]
    L.vgx_timed_launches.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int]
    _lib = L
    return L


def check(rc, what=""):
    if rc == VGX_OK:
        return
    msg = lib().vgx_strerror(rc).decode()
    if rc == VGX_ENAN:
        raise ValueError("integrand evaluates to nan")
    if rc == VGX_EINVAL:
        raise ValueError("%s: %s" % (what, msg))
    if rc == VGX_ENOMEM:
        raise MemoryError("%s: %s" % (what, msg))
    raise VgxError("%s failed: %s (%d)" % (what, msg, rc))


def torch_cuda():
    """torch with a usable gfx950 device, else a loud error."""
    import torch
    if not torch.cuda.is_available():
        raise VgxError("vegas_amd needs a ROCm GPU (MI355X / gfx950); none is visible")
    return torch


def ctx():
    """The per-device scratch context (vgx_ctx*), created on first use."""
    torch = torch_cuda()
    dev = torch.cuda.current_device()
    c = _ctx.get(dev)
    if c is None:
        p = C.c_void_p()
        check(lib().vgx_ctx_create(C.byref(p)), "vgx_ctx_create")
        c = _ctx[dev] = p
    return c


TIMED_TAGS = {1: 'fused_kernel', 2: 'b_fused_kernel', 3: 'flow_backward_kernel', 4: 'flow_wgrad_kernel',
              5: 'train_rows_kernel', 6: 'flow_sample_kernel', 7: 'cov'}


def timed_launches(cap=256):
    """The launches recorded since vgx_set_timing(ctx, 1): a list of
    (kernel name, ms) in launch order."""
    ms = (C.c_float * cap)()
    tags = (C.c_int32 * cap)()
    n = lib().vgx_timed_launches(ctx(), ms, tags, cap)
    if n < 0:
        check(n, 'vgx_timed_launches')
    return [(TIMED_TAGS.get(int(tags[i]), str(int(tags[i]))), float(ms[i])) for i in range(n)]


def stream_ptr():
    torch = torch_cuda()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def i32arr(vals):
    a = (C.c_int32 * len(vals))(*[int(v) for v in vals])
    return a


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)
