"""AdaptiveMap on the GPU -- the reference's ``vegas.AdaptiveMap``
(src/vegas/_vegas.pyx:39-832) with its state resident in HBM.

The grid lives on the device as a float64 tensor ``[dim, ld]`` (nodes of axis
d in ``[d, 0..ninc[d]]``, padding repeats the last node).  ``map``,
``invmap``, ``jac1d``, ``add_training_data`` and ``adapt`` run as HIP kernels
through libvgx; host arrays passed in are copied to the device and results
copied back into the caller's arrays, as the reference fills ``x``/``jac`` in
place.  Construction from node lists and ``make_uniform`` are host
arithmetic (``numpy.linspace``), identical to the reference.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

TINY = 10 ** (-307 + 50)


def _as_device(torch, a, dtype):
    if isinstance(a, torch.Tensor):
        if a.is_cuda:
            return a.to(dtype).contiguous(), False
        return a.to(dtype=dtype).contiguous().cuda(), True
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda(), True


class AdaptiveMap:
    r"""Adaptive map ``y -> x(y)`` (pyx:39-111).

    Args:
        grid: ``grid[d]`` is the list of nodes of axis ``d`` (at least 2),
            or another :class:`AdaptiveMap` to copy.
        ninc: increments per axis of the new grid (int or array); the new
            grid reproduces the old Jacobian (pyx:133-137).
    """

    def __init__(self, grid, ninc=None):
        torch = _lib.torch_cuda()
        if isinstance(grid, AdaptiveMap):
            self.ninc = np.array(grid.ninc)
            self._grid = grid._grid.clone()
        else:
            dim = len(grid)
            lens = np.array([len(g) for g in grid], dtype=np.intp)
            if min(lens) < 2:
                raise ValueError('grid[d] must have at least 2 elements, not {}'.format(min(lens)))
            ld = int(max(lens))
            g = np.empty((dim, ld), float)
            for d in range(dim):
                nodes = sorted(float(v) for v in grid[d])
                g[d, :len(nodes)] = nodes
                g[d, len(nodes):] = nodes[-1]
            self.ninc = lens - 1
            self._grid = torch.as_tensor(g, dtype=torch.float64).cuda()
        self.clear()
        if ninc is not None and not np.all(ninc == self.ninc):
            if np.all(np.asarray(self.ninc) == 1):
                self.make_uniform(ninc=ninc)
            else:
                self.adapt(ninc=ninc)

    # ---------------------------------------------------------- properties
    @property
    def dim(self):
        return self._grid.shape[0]

    @property
    def ld(self):
        return self._grid.shape[1]

    @property
    def grid(self):
        """Host copy of the nodes, ``[dim, max(ninc)+1]``."""
        return self._grid.cpu().numpy()

    @property
    def inc(self):
        """Host copy of the increments grid[d, i+1] - grid[d, i]."""
        g = self.grid
        inc = np.zeros((self.dim, self.ld - 1))
        for d in range(self.dim):
            n = self.ninc[d]
            inc[d, :n] = g[d, 1:n + 1] - g[d, :n]
        return inc

    @property
    def sum_f(self):
        return None if self._sum_f is None else self._sum_f.cpu().numpy()

    @property
    def n_f(self):
        return None if self._n_f is None else self._n_f.cpu().numpy()

    def region(self, d=-1):
        """x-space interval of axis d, or of all axes (pyx:144-154)."""
        if d < 0:
            return [self.region(d) for d in range(self.dim)]
        g = self.grid
        return (g[d, 0], g[d, self.ninc[d]])

    def extract_grid(self):
        """List of lists of nodes (pyx:156-163)."""
        g = self.grid
        return [list(g[d, :self.ninc[d] + 1]) for d in range(self.dim)]

    def __reduce__(self):
        return (AdaptiveMap, (self.extract_grid(),))

    def settings(self, ngrid=5):
        """Text listing of grid nodes (pyx:169-195)."""
        ans = []
        if ngrid > 0:
            g = self.grid
            for d in range(self.dim):
                grid_d = np.array(g[d, :self.ninc[d] + 1])
                nskip = int(self.ninc[d] // ngrid)
                if nskip < 1:
                    nskip = 1
                start = nskip // 2
                ans += ["    grid[%2d] = %s" % (d, np.array2string(grid_d[start::nskip], precision=3,
                                                                  prefix='    grid[xx] = '))]
        return '\n'.join(ans) + '\n'

    def clear(self):
        """Drop accumulated training data (pyx:596-599)."""
        self._sum_f = None
        self._n_f = None

    def _ninc32(self):
        return _lib.i32arr(self.ninc)

    # ---------------------------------------------------------------- maps
    def make_uniform(self, ninc=None):
        """Uniform grid with ``ninc`` increments per axis (pyx:205-241)."""
        torch = _lib.torch_cuda()
        if ninc is None:
            ninc = np.asarray(self.ninc)
        elif np.shape(ninc) == ():
            ninc = np.full(self.dim, int(ninc), dtype=np.intp)
        elif np.shape(ninc) == (self.dim,):
            ninc = np.asarray(ninc)
        else:
            raise ValueError('ninc has wrong shape -- {}'.format(np.shape(ninc)))
        if min(ninc) < 1:
            raise ValueError("no of increments < 1 in AdaptiveMap -- %s" % str(ninc))
        old = self.grid
        ld = int(max(ninc)) + 1
        g = np.empty((self.dim, ld), float)
        for d in range(self.dim):
            nodes = np.linspace(old[d, 0], old[d, self.ninc[d]], ninc[d] + 1)
            g[d, :ninc[d] + 1] = nodes
            g[d, ninc[d] + 1:] = nodes[-1]
        self.ninc = np.array(ninc, dtype=np.intp)
        self._grid = torch.as_tensor(g, dtype=torch.float64).cuda()
        self.clear()

    def map(self, y, x=None, jac=None, ny=-1):
        """Map ``y[j, d]`` to ``x[j, d]`` with Jacobian ``jac[j]`` (pyx:310-360).
        Fills ``x``/``jac`` in place when given (numpy or torch); returns
        ``(x, jac)``."""
        torch = _lib.torch_cuda()
        ny_all = y.shape[0]
        if ny < 0:
            ny = ny_all
        elif ny > ny_all:
            raise ValueError('ny > y.shape[0]: %d > %d' % (ny, ny_all))
        is_f32 = isinstance(y, torch.Tensor) and y.dtype == torch.float32
        dt = torch.float32 if is_f32 else torch.float64
        yd, _ = _as_device(torch, y[:ny], dt)
        xd = torch.empty_like(yd)
        jd = torch.empty(ny, dtype=dt, device=yd.device)
        fn = _lib.lib().vgx_map_f32 if is_f32 else _lib.lib().vgx_map_f64
        _lib.check(fn(self.dim, self._ninc32(), _lib.ptr(self._grid), self.ld, _lib.ptr(yd),
                      _lib.ptr(xd), _lib.ptr(jd), ny, _lib.stream_ptr()), 'map')
        return self._store(torch, x, xd, ny), self._store(torch, jac, jd, ny)

    @staticmethod
    def _store(torch, out, dev, n):
        if out is None:
            return dev
        if isinstance(out, torch.Tensor):
            out[:n] = dev.to(out.dtype)
            return out
        out[:n] = dev.cpu().numpy()
        return out

    def __call__(self, y):
        """x values for y (pyx:243-263); returns the same kind of array as y."""
        torch = _lib.torch_cuda()
        host = not isinstance(y, torch.Tensor)
        y = np.asarray(y, float) if host else y
        shape = y.shape
        y2 = y.reshape(-1, shape[-1])
        x, _ = self.map(y2)
        x = x.reshape(shape)
        return x.cpu().numpy() if host else x

    def jac1d(self, y):
        """Per-axis Jacobian dx[d]/dy[d] (pyx:265-295)."""
        torch = _lib.torch_cuda()
        host = not isinstance(y, torch.Tensor)
        y = np.asarray(y, float) if host else y
        shape = y.shape
        yd, _ = _as_device(torch, y.reshape(-1, shape[-1]), torch.float64)
        jd = torch.empty_like(yd)
        _lib.check(_lib.lib().vgx_jac1d_f64(self.dim, self._ninc32(), _lib.ptr(self._grid), self.ld,
                                            _lib.ptr(yd), _lib.ptr(jd), yd.shape[0],
                                            _lib.stream_ptr()), 'jac1d')
        jd = jd.reshape(shape)
        return jd.cpu().numpy() if host else jd

    def jac(self, y):
        """Jacobian dx/dy (pyx:297-306)."""
        j = self.jac1d(y)
        return np.prod(j, axis=-1) if isinstance(j, np.ndarray) else j.prod(dim=-1)

    def invmap(self, x, y=None, jac=None, nx=-1):
        """Map ``x`` back to ``y`` with Jacobian (pyx:362-416)."""
        torch = _lib.torch_cuda()
        nx_all = x.shape[0]
        if nx < 0:
            nx = nx_all
        elif nx > nx_all:
            raise ValueError('nx > x.shape[0]: %d > %d' % (nx, nx_all))
        xd, _ = _as_device(torch, x[:nx], torch.float64)
        yd = torch.empty_like(xd)
        jd = torch.empty(nx, dtype=torch.float64, device=xd.device)
        _lib.check(_lib.lib().vgx_invmap_f64(self.dim, self._ninc32(), _lib.ptr(self._grid),
                                             self.ld, _lib.ptr(xd), _lib.ptr(yd), _lib.ptr(jd), nx,
                                             _lib.stream_ptr()), 'invmap')
        return self._store(torch, y, yd, nx), self._store(torch, jac, jd, nx)

    # ----------------------------------------------------------- training
    def add_training_data(self, y, f, ny=-1, ordered=True):
        """Accumulate ``|f|`` per increment for ``adapt`` (pyx:421-464).
        ``ordered=True`` adds each bin in point order, bit-for-bit the
        reference; ``ordered=False`` uses LDS atomics (sums equal to
        rounding)."""
        torch = _lib.torch_cuda()
        if self._sum_f is None:
            shape = (self.dim, self.ld - 1)
            self._sum_f = torch.zeros(shape, dtype=torch.float64, device='cuda')
            self._n_f = torch.full(shape, TINY, dtype=torch.float64, device='cuda')
        n_all = y.shape[0]
        if ny < 0:
            ny = n_all
        elif ny > n_all:
            raise ValueError('ny > y.shape[0]: %d > %d' % (ny, n_all))
        yd, _ = _as_device(torch, y[:ny], torch.float64)
        fd, _ = _as_device(torch, f[:ny], torch.float64)
        _lib.check(_lib.lib().vgx_train_hist(_lib.ctx(), self.dim, self._ninc32(), self.ld,
                                             _lib.ptr(yd), _lib.ptr(fd), ny, _lib.ptr(self._sum_f),
                                             _lib.ptr(self._n_f), int(bool(ordered)),
                                             _lib.stream_ptr()), 'add_training_data')

    def adapt(self, alpha=0.0, ninc=None, exact=True):
        """Adapt the grid to the training data (pyx:467-594).  ``exact=True``
        follows the reference's sequential order of operations on the device;
        ``exact=False`` uses workgroup-parallel sums and CDF inversion."""
        torch = _lib.torch_cuda()
        if ninc is None:
            new_ninc = np.array(self.ninc)
        elif np.shape(ninc) == ():
            new_ninc = np.full(self.dim, int(ninc), np.intp)
        elif len(ninc) == self.dim:
            new_ninc = np.array(ninc, np.intp)
        else:
            raise ValueError('badly formed ninc = ' + str(ninc))
        if min(new_ninc) < 1:
            raise ValueError('ninc < 1: ' + str(list(new_ninc)))
        new_ld = int(max(new_ninc)) + 1
        new_grid = torch.empty((self.dim, new_ld), dtype=torch.float64, device='cuda')
        has = self._sum_f is not None
        _lib.check(_lib.lib().vgx_adapt(_lib.ctx(), self.dim, self._ninc32(), _lib.ptr(self._grid),
                                        self.ld, int(has), _lib.ptr(self._sum_f if has else None),
                                        _lib.ptr(self._n_f if has else None), float(alpha),
                                        _lib.i32arr(new_ninc), _lib.ptr(new_grid), new_ld,
                                        int(bool(exact)), _lib.stream_ptr()), 'adapt')
        self._grid = new_grid
        self.ninc = new_ninc
        self.clear()

    def adapt_to_samples(self, x, f, nitn=5, alpha=1.0, nproc=1):
        """Adapt the map to data ``{x, f(x)}`` (pyx:728-804); the inverse map,
        training histogram and adapt all run on the device."""
        torch = _lib.torch_cuda()
        x = np.ascontiguousarray(x) if not isinstance(x, torch.Tensor) else x
        if len(x.shape) != 2 or x.shape[1] != self.dim:
            raise ValueError('incompatible shape of x: {}'.format(x.shape))
        fx = f(x) if callable(f) else f
        fx = np.ascontiguousarray(fx) if not isinstance(fx, torch.Tensor) else fx
        if fx.shape[0] != x.shape[0]:
            raise ValueError('shape of x and f(x) mismatch: {} vs {}'.format(x.shape, fx.shape))
        from .integrator import Integrator  # the class default, as pyx:784 reads it
        old_ninc = int(max(max(self.ninc), Integrator.defaults['maxinc_axis']))
        tmp_ninc = int(min(old_ninc, x.shape[0] / 10.))
        if tmp_ninc < 2:
            raise ValueError('not enough samples: {}'.format(x.shape[0]))
        xd, _ = _as_device(torch, x, torch.float64)
        fd, _ = _as_device(torch, fx, torch.float64)
        self.adapt(ninc=tmp_ninc)
        for _ in range(nitn):
            yd, jd = self.invmap(xd)
            self.add_training_data(yd, (jd * fd) ** 2)
            self.adapt(alpha=alpha, ninc=tmp_ninc)
        if np.any(tmp_ninc != old_ninc):
            self.adapt(ninc=old_ninc)

    def random(self, n=None, rng=None):
        """Random points in x space drawn uniformly in y (pyx:197-203)."""
        rng = np.random.default_rng() if rng is None else rng
        y = rng.random(self.dim if n is None else (n, self.dim))
        return self(y)
