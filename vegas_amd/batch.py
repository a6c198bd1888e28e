"""Many independent vegas+ integrals advanced together on one GPU.

:class:`BatchIntegrator` runs n integrals of single-Gaussian integrands
``f_i(x) = A_i exp(-B_i |x - c_i|^2)`` on ``[0,1]^dim_i`` (dim_i <= 10), each
with the geometry :class:`vegas_amd.Integrator` would give it (same
``nstrat``, ``ninc``, ``min_neval_hcube``; ``maxinc_axis`` 500 by default)
and its own Philox stream, so integral i sees exactly the samples of a
separate ``Integrator(seed=seeds[i])`` run.  All n advance in one stream of
launches per iteration (libvgx ``vgx_batch_*``): one allocation, one
persistent sampler launch pulling work items (runs of whole wave tiles of
one integral) from a queue, one pass over cut hypercubes, one close and one
adapt.  The iterations of a call are captured as one hipGraph and replayed
by later calls.  The work-item size is chosen by :meth:`BatchIntegrator.autotune`.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import zlib

import numpy as np

from . import _geometry, _lib
from ._ravg import gammaQ

MAXDIM = 10
MAXINC = 500
LD = MAXINC + 1


class _BatchConfig(C.Structure):
    _fields_ = [("n", C.c_int32), ("dim", C.c_void_p), ("ninc", C.c_void_p), ("nstrat", C.c_void_p),
                ("neval", C.c_void_p), ("min_neval_hcube", C.c_void_p), ("max_neval_hcube", C.c_void_p),
                ("seed", C.c_void_p), ("neval_frac", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("tiles_per_item", C.c_int32), ("blocks", C.c_int32)]


class _BatchState(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("sigf", C.c_void_p), ("sum_sigf", C.c_void_p),
                ("itn_mean", C.c_void_p), ("itn_var", C.c_void_p), ("itn_neval", C.c_void_p),
                ("ravg", C.c_void_p), ("status", C.c_void_p)]


def _bind(L):
    if getattr(L, "_vgx_batch_bound", False):
        return L
    vp = C.c_void_p
    L.vgx_batch_create.argtypes = [vp, C.POINTER(_BatchConfig), C.POINTER(vp)]
    L.vgx_batch_destroy.argtypes = [vp]
    L.vgx_batch_sizes.argtypes = [vp, C.POINTER(C.c_int64), vp]
    L.vgx_batch_integrate.argtypes = [vp, vp, vp, vp, C.POINTER(_BatchState), C.c_int, vp]
    L.vgx_batch_set_iteration.argtypes = [vp, C.c_uint32]
    L._vgx_batch_bound = True
    return L


class BatchResult:
    """Per-integral results of one call: running averages over its nitn
    iterations (weighted, as RAvg) and the per-iteration estimates."""

    def __init__(self, itn_mean, itn_var, itn_neval, ravg):
        self.itn_mean = itn_mean          # [nitn, n]
        self.itn_sdev = np.sqrt(itn_var)  # [nitn, n]
        self.itn_neval = itn_neval        # [nitn, n]
        self.mean = ravg[:, 0]
        self.sdev = ravg[:, 1]
        self.chi2 = ravg[:, 2]
        self.dof = ravg[:, 3].astype(int)
        self.Q = np.array([gammaQ(d / 2., c / 2.) if d > 0 else 1.0 for d, c in zip(self.dof, self.chi2)])
        self.sum_neval = int(itn_neval.sum())

    def __len__(self):
        return len(self.mean)


class BatchIntegrator:
    """n independent vegas+ integrals of single Gaussians (see module doc).

    dims, nevals: per-integral dimension (1-10) and evaluations per
    iteration; seeds: per-integral Philox keys (default 1..n).  alpha, beta,
    neval_frac, max_neval_hcube, maxinc_axis: as for Integrator, shared.
    grids: optional initial grids, a list of (dim_i, ninc_i + 1) arrays.
    tiles_per_item: work-item size in 256-point wave tiles (1-240); None
    takes the autotuned value when one is cached, else 128.
    """

    def __init__(self, dims, nevals, seeds=None, alpha=0.5, beta=0.75, neval_frac=0.75,
                 max_neval_hcube=50000, maxinc_axis=MAXINC, tiles_per_item=None, grids=None,
                 blocks=0):
        torch = _lib.torch_cuda()
        dims = np.asarray(dims, np.int64)
        nevals = np.asarray(nevals, np.int64)
        n = len(dims)
        if n < 1 or len(nevals) != n:
            raise ValueError('dims and nevals need one entry per integral')
        if np.any(dims < 1) or np.any(dims > MAXDIM):
            raise ValueError('dimensions must be 1..%d' % MAXDIM)
        if not beta > 0:
            raise ValueError('the batched integrator needs beta > 0')
        if maxinc_axis > MAXINC:
            raise ValueError('maxinc_axis <= %d' % MAXINC)
        self.n = n
        self.dims = dims
        self.alpha, self.beta, self.neval_frac = float(alpha), float(beta), float(neval_frac)
        seeds = np.arange(1, n + 1, dtype=np.uint64) if seeds is None else np.asarray(seeds, np.uint64)
        ninc = np.zeros((n, MAXDIM), np.int32)
        nstrat = np.zeros((n, MAXDIM), np.int32)
        nmin = np.zeros(n, np.int64)
        self.nhcube = np.zeros(n, np.int64)
        self.neval = np.zeros(n, np.int64)
        for i in range(n):
            g = _geometry.compute(int(dims[i]), neval=int(nevals[i]), neval_frac=neval_frac, beta=beta,
                                  maxinc_axis=maxinc_axis, max_neval_hcube=max_neval_hcube)
            d = int(dims[i])
            ninc[i, :d] = g['ninc']
            nstrat[i, :d] = g['nstrat']
            nmin[i] = g['min_neval_hcube']
            self.nhcube[i] = g['nhcube']
            self.neval[i] = g['neval']
        nmax = np.full(n, int(max_neval_hcube), np.int64)
        self.ninc, self.nstrat, self.min_neval_hcube = ninc, nstrat, nmin
        self.seeds = seeds
        self._host = (dims.astype(np.int32), ninc, nstrat, self.neval, nmin, nmax, seeds)
        if tiles_per_item is None:
            tiles_per_item = _tuned(self._signature()) or 128
        self.tiles_per_item = int(tiles_per_item)
        self.blocks = int(blocks)
        self._L = _bind(_lib.lib())
        self._h = self._create(self.tiles_per_item, self.blocks)
        tot = C.c_int64()
        base = np.zeros(n, np.int64)
        _lib.check(self._L.vgx_batch_sizes(self._h, C.byref(tot), base.ctypes.data), 'vgx_batch_sizes')
        self.hcube_base = base
        dev = torch.device('cuda', torch.cuda.current_device())
        self._torch = torch
        self.grid = torch.empty((n, MAXDIM, LD), dtype=torch.float64, device=dev)
        self.sigf = torch.ones(int(tot.value), dtype=torch.float64, device=dev)
        self.sum_sigf = torch.as_tensor(self.nhcube.astype(np.float64), device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self._set_grids(grids)
        self._bufs = {}

    # -------------------------------------------------------------- helpers
    def _signature(self):
        h = np.concatenate([self.dims, self.neval]).astype(np.int64)
        return '%d:%d:%08x' % (self.n, int(self.neval.sum()), zlib.crc32(h.tobytes()))

    def _create(self, tpi, blocks):
        dims, ninc, nstrat, neval, nmin, nmax, seeds = self._host
        cfg = _BatchConfig()
        cfg.n = self.n
        cfg.dim, cfg.ninc, cfg.nstrat = dims.ctypes.data, ninc.ctypes.data, nstrat.ctypes.data
        cfg.neval, cfg.min_neval_hcube, cfg.max_neval_hcube = neval.ctypes.data, nmin.ctypes.data, nmax.ctypes.data
        cfg.seed = seeds.ctypes.data
        cfg.neval_frac, cfg.alpha, cfg.beta = self.neval_frac, self.alpha, self.beta
        cfg.tiles_per_item, cfg.blocks = int(tpi), int(blocks)
        h = C.c_void_p()
        _lib.check(self._L.vgx_batch_create(_lib.ctx(), C.byref(cfg), C.byref(h)), 'vgx_batch_create')
        return h

    def _set_grids(self, grids):
        g = np.zeros((self.n, MAXDIM, LD))
        for i in range(self.n):
            for d in range(int(self.dims[i])):
                k = int(self.ninc[i, d])
                if grids is not None:
                    row = np.asarray(grids[i][d], np.float64)[:k + 1]
                else:
                    row = np.linspace(0.0, 1.0, k + 1)
                g[i, d, :k + 1] = row
                g[i, d, k + 1:] = row[-1]
        self.grid.copy_(self._torch.as_tensor(g))

    def map_grid(self, i):
        """Grid of integral i as a (dim, ninc+1) numpy array per axis list."""
        g = self.grid[i].cpu().numpy()
        return [g[d, :int(self.ninc[i, d]) + 1] for d in range(int(self.dims[i]))]

    def _state(self, nitn):
        torch = self._torch
        key = nitn
        if key not in self._bufs:
            dev = self.grid.device
            self._bufs[key] = (torch.zeros((nitn, self.n), dtype=torch.float64, device=dev),
                               torch.zeros((nitn, self.n), dtype=torch.float64, device=dev),
                               torch.zeros((nitn, self.n), dtype=torch.int64, device=dev),
                               torch.zeros((self.n, 4), dtype=torch.float64, device=dev))
        m, v, ne, ra = self._bufs[key]
        st = _BatchState()
        st.grid, st.sigf, st.sum_sigf = self.grid.data_ptr(), self.sigf.data_ptr(), self.sum_sigf.data_ptr()
        st.itn_mean, st.itn_var, st.itn_neval = m.data_ptr(), v.data_ptr(), ne.data_ptr()
        st.ravg, st.status = ra.data_ptr(), self.status.data_ptr()
        return st, (m, v, ne, ra)

    def _params(self, A, B, centers):
        torch = self._torch
        dev = self.grid.device
        A = np.array(np.broadcast_to(np.asarray(A, np.float64), (self.n,)))
        B = np.array(np.broadcast_to(np.asarray(B, np.float64), (self.n,)))
        c = np.zeros((self.n, MAXDIM))
        for i, ci in enumerate(centers):
            ci = np.asarray(ci, np.float64)
            c[i, :len(ci)] = ci
        key = (A.tobytes(), B.tobytes(), c.tobytes())
        if getattr(self, '_pkey', None) != key:
            self._pdev = tuple(torch.as_tensor(np.ascontiguousarray(a), device=dev) for a in (A, B, c))
            self._pkey = key
        return self._pdev

    # ------------------------------------------------------------------ API
    def launch(self, A, B, centers, nitn=8):
        """Enqueue nitn iterations on the current stream; returns the device
        result buffers (itn_mean, itn_var, itn_neval, ravg) without waiting."""
        pA, pB, pc = self._params(A, B, centers)
        st, bufs = self._state(int(nitn))
        _lib.check(self._L.vgx_batch_integrate(self._h, C.c_void_p(pA.data_ptr()), C.c_void_p(pB.data_ptr()),
                                               C.c_void_p(pc.data_ptr()), C.byref(st), int(nitn),
                                               _lib.stream_ptr()), 'vgx_batch_integrate')
        self._keep = (pA, pB, pc)
        return bufs

    def __call__(self, A, B, centers, nitn=8):
        m, v, ne, ra = self.launch(A, B, centers, nitn)
        self._torch.cuda.current_stream().synchronize()
        s = int(self.status.item())
        if s == _lib.VGX_ENAN:
            raise ValueError('integrand evaluates to nan')
        if s != 0:
            _lib.check(s, 'vgx_batch_integrate')
        return BatchResult(m.cpu().numpy(), v.cpu().numpy(), ne.cpu().numpy(), ra.cpu().numpy())

    def set_iteration(self, it):
        _lib.check(self._L.vgx_batch_set_iteration(self._h, int(it)), 'vgx_batch_set_iteration')

    def autotune(self, A, B, centers, candidates=(64, 128, 192, 240), nitn=2, save=True):
        """Time nitn iterations for each work-item size on a copy of the
        current state and keep the fastest (cached per workload signature).
        The samples never depend on the item size; only the order of the
        compensated partial sums does."""
        torch = self._torch
        saved = (self.grid.clone(), self.sigf.clone(), self.sum_sigf.clone())
        times = {}
        h0, t0 = self._h, self.tiles_per_item
        for tpi in candidates:
            h = self._create(tpi, self.blocks)
            self._h = h
            best = float('inf')
            for rep in range(2):
                self.grid.copy_(saved[0])
                self.sigf.copy_(saved[1])
                self.sum_sigf.copy_(saved[2])
                self.set_iteration(0)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                self.launch(A, B, centers, nitn)
                e1.record()
                e1.synchronize()
                if rep:  # the first run includes graph capture
                    best = min(best, e0.elapsed_time(e1))
            times[int(tpi)] = best
            self._L.vgx_batch_destroy(h)
        self._h = h0
        self.grid.copy_(saved[0])
        self.sigf.copy_(saved[1])
        self.sum_sigf.copy_(saved[2])
        self.set_iteration(0)
        bt = min(times, key=times.get)
        if bt != t0:
            self._L.vgx_batch_destroy(self._h)
            self._h = self._create(bt, self.blocks)
            self.tiles_per_item = bt
        if save:
            _store_tuned(self._signature(), bt)
        self.tune_times = times
        return bt

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                self._L.vgx_batch_destroy(h)
            except Exception:
                pass


def _cache_path():
    return os.environ.get('VEGAS_AMD_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'vegas_amd', 'batch_tune.json'))


def _tuned(sig):
    try:
        with open(_cache_path()) as fh:
            return json.load(fh).get(sig)
    except (OSError, ValueError):
        return None


def _store_tuned(sig, tpi):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        d = {}
        if os.path.exists(p):
            with open(p) as fh:
                d = json.load(fh)
        d[sig] = int(tpi)
        with open(p, 'w') as fh:
            json.dump(d, fh)
    except (OSError, ValueError):
        pass


def gaussian_params(dims, sigmas):
    """A, B of normalised isotropic Gaussians: A = (2 pi s^2)^(-d/2), B = 1/(2 s^2)."""
    A = np.array([(2 * math.pi * s * s) ** (-0.5 * d) for d, s in zip(dims, sigmas)])
    B = np.array([0.5 / (s * s) for s in sigmas])
    return A, B
