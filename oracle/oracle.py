"""ctypes front end of the FP64 CPU oracle (oracle/vegas_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this module, and only as the
checker or the timed CPU baseline.  ``vegas_amd`` never imports it.

Each wrapper restates one reference routine of gplepage/vegas 6.4.1
(``src/vegas/_vegas.pyx``, cited ``pyx:N``); the arithmetic is in the C
library, compiled with ``-ffp-contract=off``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

FCN_T = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_int,
                    C.POINTER(C.c_double))


class _GaussMix(C.Structure):
    _fields_ = [("dim", C.c_int), ("ncomp", C.c_int), ("A", C.POINTER(C.c_double)),
                ("B", C.POINTER(C.c_double)), ("c", C.POINTER(C.c_double))]


class _GeomOut(C.Structure):
    _fields_ = [("nstrat", C.c_int64 * 64), ("ninc", C.c_int64 * 64), ("nhcube", C.c_int64),
                ("neval", C.c_int64), ("min_neval_hcube", C.c_int64),
                ("workspace", C.c_int64), ("status", C.c_int)]


class _IterParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("ld", C.c_int), ("nstrat", C.POINTER(C.c_int64)),
                ("nhcube", C.c_int64), ("neval", C.c_int64), ("neval_frac", C.c_double),
                ("beta", C.c_double), ("alpha", C.c_double), ("adapt", C.c_int),
                ("adapt_to_errors", C.c_int), ("min_neval_hcube", C.c_int64),
                ("max_neval_hcube", C.c_int64), ("min_neval_batch", C.c_int64),
                ("seed", C.c_uint64), ("iteration", C.c_uint32), ("rng_f32", C.c_int)]


class _IterOut(C.Structure):
    _fields_ = [("mean", C.c_double), ("var", C.c_double), ("sum_sigf", C.c_double),
                ("last_neval", C.c_int64), ("nh_min", C.c_int64), ("nh_max", C.c_int64),
                ("nbatch", C.c_int64)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.vgo_tiny.restype = C.c_double
        L.vgo_huge.restype = C.c_double
        L.vgo_epsilon.restype = C.c_double
        L.vgo_philox4x32_10.argtypes = [_up, _up, _up]
        for nm in ("vgo_uniforms_f64", "vgo_uniforms_f32"):
            getattr(L, nm).argtypes = [C.c_uint64, C.c_uint32, C.c_int64, C.c_int64, C.c_int, _dp]
        L.vgo_make_uniform.argtypes = [C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp]
        L.vgo_map.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, C.c_int64]
        L.vgo_invmap.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, C.c_int64]
        L.vgo_jac1d.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp, _dp, C.c_int64]
        L.vgo_add_training_data.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_int64, _dp, _dp]
        L.vgo_adapt.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, C.c_double,
                                _ip, C.c_int, _dp, _dp]
        L.vgo_geometry.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_void_p, C.c_double,
                                   C.c_double, C.c_int, C.c_int, C.c_int64, C.c_double, C.c_int,
                                   C.c_int64, C.c_int64, C.POINTER(_GeomOut)]
        L.vgo_gaussmix_eval.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_int,
                                        C.POINTER(C.c_double)]
        L.vgo_iterate.argtypes = [C.POINTER(_IterParams), _dp, _dp, _ip, _dp, C.c_double, _dp,
                                  _dp, C.c_void_p, C.c_void_p, C.POINTER(_IterOut)]
        L.vgo_run.argtypes = [C.POINTER(_IterParams), C.c_int, _dp, _dp, _ip, _dp,
                              C.POINTER(C.c_double), C.c_void_p, C.c_void_p, _dp, _dp, _ip]
        L.vgo_ravg.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


TINY = 1e-257
EPSILON = np.finfo(float).eps * 1e4


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().vgo_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def uniforms(seed, iteration, p0, n, dim, f32=False):
    U = np.empty((n, dim), np.float64)
    fn = lib().vgo_uniforms_f32 if f32 else lib().vgo_uniforms_f64
    fn(seed, iteration, p0, n, dim, U)
    return U


class OracleMap:
    """AdaptiveMap state (pyx:112-137): grid dim x ld, inc dim x (ld-1), ninc.
    Padding beyond ninc repeats the last node (inc padding 0)."""

    def __init__(self, grid, ninc=None):
        if isinstance(grid, OracleMap):
            self.grid = grid.grid.copy(); self.inc = grid.inc.copy(); self.ninc = grid.ninc.copy()
        else:
            dim = len(grid)
            lens = [len(g) for g in grid]
            if min(lens) < 2:
                raise ValueError("grid[d] must have at least 2 elements")
            ld = max(lens)
            self.ninc = np.array(lens, np.int64) - 1
            self.grid = np.empty((dim, ld), np.float64)
            self.inc = np.zeros((dim, ld - 1), np.float64)
            for d in range(dim):
                g = sorted(float(v) for v in grid[d])
                self.grid[d, :len(g)] = g
                self.grid[d, len(g):] = g[-1]
                for i in range(len(g) - 1):
                    self.inc[d, i] = self.grid[d, i + 1] - self.grid[d, i]
        self.clear()
        if ninc is not None:
            nn = np.full(self.dim, int(ninc), np.int64) if np.shape(ninc) == () else \
                np.asarray(ninc, np.int64)
            if not np.all(nn == self.ninc):
                if np.all(self.ninc == 1):
                    self.make_uniform(nn)
                else:
                    self.adapt(ninc=nn)

    @property
    def dim(self):
        return self.grid.shape[0]

    @property
    def ld(self):
        return self.grid.shape[1]

    def clear(self):
        self.sum_f = None
        self.n_f = None

    def make_uniform(self, ninc):
        ninc = np.asarray(ninc, np.int64)
        lo = np.ascontiguousarray(self.grid[:, 0])
        hi = np.array([self.grid[d, self.ninc[d]] for d in range(self.dim)])
        ld = int(ninc.max()) + 1
        g = np.empty((self.dim, ld)); inc = np.empty((self.dim, ld - 1))
        lib().vgo_make_uniform(self.dim, lo, hi, ninc, ld, g, inc)
        self.grid, self.inc, self.ninc = g, inc, ninc.copy()
        self.clear()

    def map(self, y):
        y = np.ascontiguousarray(y, np.float64)
        x = np.empty_like(y); jac = np.empty(y.shape[0])
        lib().vgo_map(self.dim, self.ld, self.grid, self.inc, self.ninc, y, x, jac, y.shape[0])
        return x, jac

    def invmap(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x); jac = np.empty(x.shape[0])
        lib().vgo_invmap(self.dim, self.ld, self.grid, self.inc, self.ninc, x, y, jac, x.shape[0])
        return y, jac

    def jac1d(self, y):
        y = np.ascontiguousarray(y, np.float64)
        j = np.empty_like(y)
        lib().vgo_jac1d(self.dim, self.ld, self.inc, self.ninc, y, j, y.shape[0])
        return j

    def add_training_data(self, y, f):
        y = np.ascontiguousarray(y, np.float64)
        f = np.ascontiguousarray(f, np.float64)
        if self.sum_f is None:
            self.sum_f = np.zeros(self.inc.shape)
            self.n_f = np.zeros(self.inc.shape) + TINY
        lib().vgo_add_training_data(self.dim, self.ld, self.ninc, y, f, y.shape[0], self.sum_f,
                                    self.n_f)

    def adapt(self, alpha=0.0, ninc=None):
        if ninc is None:
            nn = self.ninc.copy()
        elif np.shape(ninc) == ():
            nn = np.full(self.dim, int(ninc), np.int64)
        else:
            nn = np.asarray(ninc, np.int64)
        if nn.min() < 1:
            raise ValueError("ninc < 1")
        new_ld = int(nn.max()) + 1
        g = np.empty((self.dim, new_ld)); inc = np.zeros((self.dim, new_ld - 1))
        has = self.sum_f is not None
        sf = self.sum_f if has else np.zeros(self.inc.shape)
        nf = self.n_f if has else np.zeros(self.inc.shape)
        rc = lib().vgo_adapt(self.dim, self.ld, self.grid, self.inc, self.ninc, int(has), sf, nf,
                             float(alpha), nn, new_ld, g, inc)
        if rc:
            raise ValueError("bad ninc")
        self.grid, self.inc, self.ninc = g, inc, nn
        self.clear()


def geometry(dim, neval=None, nstrat=None, neval_frac=0.75, beta=0.75, adapt_to_errors=False,
             uniform_nstrat=False, maxinc_axis=1000, max_mem=1e9, minimize_mem=False,
             min_neval_batch=100000, max_neval_hcube=50000):
    out = _GeomOut()
    ns = None
    if nstrat is not None:
        ns = np.ascontiguousarray(nstrat, np.int64)
    lib().vgo_geometry(dim, int(neval) if neval is not None else 0, int(neval is not None),
                       ns.ctypes.data if ns is not None else None, neval_frac, beta,
                       int(adapt_to_errors), int(uniform_nstrat), int(maxinc_axis),
                       float(max_mem), int(minimize_mem), int(min_neval_batch),
                       int(max_neval_hcube), C.byref(out))
    if out.status == -1:
        raise ValueError("bad geometry")
    if out.status == -2:
        raise MemoryError("work arrays larger than max_mem")
    return dict(nstrat=np.array(out.nstrat[:dim]), ninc=np.array(out.ninc[:dim]),
                nhcube=out.nhcube, neval=out.neval, min_neval_hcube=out.min_neval_hcube,
                workspace=out.workspace)


class GaussMix:
    """f(x) = sum_k A_k exp(-B_k |x - c_k|^2) evaluated in FP64 by the oracle."""

    def __init__(self, A, B, c):
        self.A = np.ascontiguousarray(A, np.float64)
        self.B = np.ascontiguousarray(B, np.float64)
        self.c = np.ascontiguousarray(c, np.float64)
        self.ncomp, self.dim = self.c.shape
        self._s = _GaussMix(self.dim, self.ncomp,
                            self.A.ctypes.data_as(C.POINTER(C.c_double)),
                            self.B.ctypes.data_as(C.POINTER(C.c_double)),
                            self.c.ctypes.data_as(C.POINTER(C.c_double)))

    def fcn_ctx(self):
        return C.cast(lib().vgo_gaussmix_eval, C.c_void_p), C.cast(C.pointer(self._s), C.c_void_p)

    def __call__(self, x):
        x = np.ascontiguousarray(x, np.float64)
        f = np.empty(x.shape[0])
        lib().vgo_gaussmix_eval(C.cast(C.pointer(self._s), C.c_void_p),
                                x.ctypes.data_as(C.POINTER(C.c_double)), x.shape[0], self.dim,
                                f.ctypes.data_as(C.POINTER(C.c_double)))
        return f


def _python_fcn(f):
    """Wrap an lbatch numpy integrand f(x[n, dim]) -> f[n] as a C callback."""

    def cb(ctx, xp, n, dim, fp):
        try:
            x = np.ctypeslib.as_array(xp, shape=(n, dim)).copy()
            fv = np.asarray(f(x), np.float64).reshape(n)
            np.ctypeslib.as_array(fp, shape=(n,))[:] = fv
            return 0
        except Exception:  # noqa: BLE001 - surfaced as a status code
            return -100

    c = FCN_T(cb)
    return c, C.cast(c, C.c_void_p)


class OracleIntegrator:
    """Reference-structured vegas+ driver (pyx:1256-1446 geometry, pyx:2086-2227
    iteration loop) with the RNG seam (pyx:1676-1680) fed by Philox."""

    def __init__(self, region, neval=1000, nstrat=None, neval_frac=0.75, alpha=0.5, beta=0.75,
                 adapt=True, adapt_to_errors=False, maxinc_axis=1000, min_neval_batch=100000,
                 max_neval_hcube=50000, seed=2026, rng_f32=False, max_mem=1e9):
        self.dim = len(region)
        g = geometry(self.dim, neval=neval, nstrat=nstrat, neval_frac=neval_frac, beta=beta,
                     adapt_to_errors=adapt_to_errors, maxinc_axis=maxinc_axis,
                     min_neval_batch=min_neval_batch, max_neval_hcube=max_neval_hcube,
                     max_mem=max_mem)
        self.__dict__.update(g)
        self.map = OracleMap([list(r) for r in region])
        # Integrator.set regrids with map.adapt(ninc=...) (pyx:1385-1386)
        if not np.all(self.map.ninc == g["ninc"]):
            self.map.adapt(ninc=g["ninc"])
        self.neval_frac, self.alpha, self.beta = neval_frac, alpha, beta
        self.adapt_, self.adapt_to_errors = adapt, adapt_to_errors
        self.min_neval_batch, self.max_neval_hcube = min_neval_batch, max_neval_hcube
        self.seed, self.rng_f32 = seed, rng_f32
        self.sigf = np.ones(self.nhcube)
        self.sum_sigf = float(self.nhcube)
        self.iteration = 0

    def _params(self, ld):
        self._ns = np.ascontiguousarray(self.nstrat, np.int64)
        return _IterParams(self.dim, ld, self._ns.ctypes.data_as(C.POINTER(C.c_int64)),
                           self.nhcube, self.neval, self.neval_frac, self.beta, self.alpha,
                           int(self.adapt_), int(self.adapt_to_errors), self.min_neval_hcube,
                           self.max_neval_hcube, self.min_neval_batch, self.seed,
                           self.iteration, int(self.rng_f32))

    def _fcn(self, f):
        if isinstance(f, GaussMix):
            return f.fcn_ctx()
        keep, ptr = _python_fcn(f)
        self._keep = keep
        return ptr, None

    def iterate(self, f, sum_f=None, n_f=None):
        """One iteration; returns (mean, var, out-struct, sum_f, n_f)."""
        m = self.map
        if sum_f is None:
            sum_f = np.zeros(m.inc.shape)
            n_f = np.zeros(m.inc.shape) + TINY
        fp, ctx = self._fcn(f)
        o = _IterOut()
        P = self._params(m.ld)
        rc = lib().vgo_iterate(C.byref(P), m.grid, m.inc, m.ninc, self.sigf, self.sum_sigf,
                               sum_f, n_f, fp, ctx, C.byref(o))
        if rc == -3:
            raise ValueError("integrand evaluates to nan")
        if rc:
            raise RuntimeError("integrand failed (%d)" % rc)
        return o, sum_f, n_f

    def run(self, f, nitn):
        """nitn iterations of Integrator.__call__; returns per-iteration arrays."""
        m = self.map
        fp, ctx = self._fcn(f)
        means = np.zeros(nitn); vars_ = np.zeros(nitn); nev = np.zeros(nitn, np.int64)
        ss = C.c_double(self.sum_sigf)
        P = self._params(m.ld)
        rc = lib().vgo_run(C.byref(P), nitn, m.grid, m.inc, m.ninc, self.sigf, C.byref(ss), fp,
                           ctx, means, vars_, nev)
        if rc == -3:
            raise ValueError("integrand evaluates to nan")
        if rc:
            raise RuntimeError("integrand failed (%d)" % rc)
        self.sum_sigf = ss.value
        self.iteration += nitn
        return means, vars_, nev


def ravg(means, vars_, weighted=True):
    out = np.zeros(3)
    lib().vgo_ravg(int(weighted), len(means), np.ascontiguousarray(means, np.float64),
                   np.ascontiguousarray(vars_, np.float64), out)
    return out


def gammaQ(a, x):
    """Regularised upper incomplete gamma Q(a, x) (series for x < a+1, Lentz
    continued fraction otherwise; Press et al., Numerical Recipes 6.2)."""
    if x < 0 or a <= 0:
        return float("nan")
    if x == 0:
        return 1.0
    gln = math.lgamma(a)
    if x < a + 1.0:
        ap, s, d = a, 1.0 / a, 1.0 / a
        for _ in range(10000):
            ap += 1.0
            d *= x / ap
            s += d
            if abs(d) < abs(s) * 1e-16:
                break
        return 1.0 - s * math.exp(-x + a * math.log(x) - gln)
    b = x + 1.0 - a
    c = 1.0 / 1e-300
    d = 1.0 / b
    h = d
    for i in range(1, 10000):
        an = -i * (i - a)
        b += 2.0
        d = an * d + b
        if abs(d) < 1e-300:
            d = 1e-300
        c = b + an / c
        if abs(c) < 1e-300:
            c = 1e-300
        d = 1.0 / d
        de = d * c
        h *= de
        if abs(de - 1.0) < 1e-16:
            break
    return math.exp(-x + a * math.log(x) - gln) * h
