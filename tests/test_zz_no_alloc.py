"""vgx_train_hist_noalloc: the ordered training histogram with its
temporaries in the library's static device scratch (no allocation), against
vgx_train_hist(ordered=1), which allocates them in a context: bit for bit,
on two calls in a row (the second smaller, so it runs on what the first
left in the scratch), and a size beyond the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from vegas_amd import _lib
    L = _lib.lib()
    L.vgx_train_hist_noalloc.argtypes = [C.c_int, C.POINTER(C.c_int32), C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    return torch, _lib, L


def _hist(torch, _lib, L, y, f, ninc, ld, noalloc):
    dim = y.shape[1]
    nb = ld - 1
    sum_f = torch.zeros((dim, nb), dtype=torch.float64, device='cuda')
    n_f = torch.full((dim, nb), 1e-257, dtype=torch.float64, device='cuda')
    yd = torch.as_tensor(y, device='cuda')
    fd = torch.as_tensor(f, device='cuda')
    ni = _lib.i32arr(ninc)
    if noalloc:
        rc = L.vgx_train_hist_noalloc(dim, ni, ld, _lib.ptr(yd), _lib.ptr(fd), y.shape[0], _lib.ptr(sum_f),
                                      _lib.ptr(n_f), _lib.stream_ptr())
    else:
        rc = L.vgx_train_hist(_lib.ctx(), dim, ni, ld, _lib.ptr(yd), _lib.ptr(fd), y.shape[0], _lib.ptr(sum_f),
                              _lib.ptr(n_f), 1, _lib.stream_ptr())
    _lib.check(rc, 'train_hist')
    torch.cuda.synchronize()
    return sum_f.cpu().numpy(), n_f.cpu().numpy()


def test_noalloc_equals_ordered_twice(env):
    torch, _lib, L = env
    rng = np.random.default_rng(7)
    ninc = [997, 500, 64, 1000, 3]
    ld = 1001
    for n in (300000, 70001):  # the second call is smaller
        y = rng.uniform(0, 1, (n, len(ninc)))
        y[:7] = 0.0
        y[7:9] = 1.0
        f = rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))
        a = _hist(torch, _lib, L, y, f, ninc, ld, noalloc=True)
        b = _hist(torch, _lib, L, y, f, ninc, ld, noalloc=False)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert a[1][0, :997].sum() > 0.99 * n


def test_noalloc_refuses_beyond_scratch(env):
    torch, _lib, L = env
    # 8 axes of 1000 increments: about 95 B of scratch per point, so 2e6
    # points need more than the 128 MiB scratch (refused before any launch)
    y = torch.zeros((1, 8), dtype=torch.float64, device='cuda')
    rc = L.vgx_train_hist_noalloc(8, _lib.i32arr([1000] * 8), 1001, _lib.ptr(y), _lib.ptr(y), 2000000,
                                  _lib.ptr(y), _lib.ptr(y), _lib.stream_ptr())
    assert rc == _lib.VGX_ERANGE
