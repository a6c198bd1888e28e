"""Running averages of iteration results and a minimal Gaussian value type.

Host logic restating the scalar ``RAvg`` of gplepage/vegas 6.4.1
(src/vegas/_vegas.pyx:2276-2451).  The reference builds on ``gvar.GVar``;
here :class:`GVar` carries only mean and standard deviation, with gvar's
compact string form ``1.35(86)``.  Inside ``Integrator.__call__`` the same
weighted average is also computed in FP64 on the device; this class is the
host view of it and the stand-alone API.
"""
from __future__ import annotations

import math

import numpy as np

TINY = 10 ** (-307 + 50)  # pyx:34, sys.float_info.min_10_exp + 50
EPSILON = 2.220446049250313e-16 * 1e4  # pyx:36, sys.float_info.epsilon * 1e4


def gammaQ(a, x):
    """Regularised upper incomplete gamma function Q(a, x) (the reference's
    gvar.gammaQ, pyx:2374)."""
    try:
        from scipy.special import gammaincc
        return float(gammaincc(a, x))
    except ImportError:  # pragma: no cover - scipy is present in this image
        return _gammaQ(a, x)


def _gammaQ(a, x):
    """Series for x < a+1, Lentz continued fraction otherwise (Press et al.,
    Numerical Recipes, section 6.2)."""
    if x <= 0:
        return 1.0
    gln = math.lgamma(a)
    if x < a + 1.0:
        ap, s, d = a, 1.0 / a, 1.0 / a
        for _ in range(100000):
            ap += 1.0
            d *= x / ap
            s += d
            if abs(d) < abs(s) * 1e-16:
                break
        return 1.0 - s * math.exp(-x + a * math.log(x) - gln)
    b = x + 1.0 - a
    c, d = 1e300, 1.0 / b
    h = d
    for i in range(1, 100000):
        an = -i * (i - a)
        b += 2.0
        d = an * d + b
        d = 1e-300 if abs(d) < 1e-300 else d
        c = b + an / c
        c = 1e-300 if abs(c) < 1e-300 else c
        d = 1.0 / d
        h *= d * c
        if abs(d * c - 1.0) < 1e-16:
            break
    return math.exp(-x + a * math.log(x) - gln) * h


def fmt_gvar(mean, sdev):
    """gvar-style compact form: two significant digits of the error."""
    if not (math.isfinite(mean) and math.isfinite(sdev)):
        return '%g +- %g' % (mean, sdev)
    if sdev <= 0:
        return '%g(0)' % mean
    ndec = 1 - int(math.floor(math.log10(sdev)))
    if ndec > 0:
        if sdev >= 1.0:
            return '%.*f(%.*f)' % (ndec, mean, ndec, sdev)
        return '%.*f(%d)' % (ndec, mean, int(round(sdev * 10 ** ndec)))
    if ndec == 0:
        return '%.0f(%d)' % (mean, int(round(sdev)))
    # error >= 100: use scientific notation like gvar (mantissa(err)e+NN)
    exp = int(math.floor(math.log10(abs(mean)))) if mean != 0 else 0
    scale = 10.0 ** exp
    return fmt_gvar(mean / scale, sdev / scale) + 'e%+03d' % exp


class GVar:
    """Gaussian random variable: mean and standard deviation."""

    def __init__(self, mean=0.0, sdev=0.0):
        self._mean = float(mean)
        self._sdev = float(sdev)

    @property
    def mean(self):
        return self._mean

    @property
    def sdev(self):
        return self._sdev

    @property
    def var(self):
        return self._sdev ** 2

    def __str__(self):
        return fmt_gvar(self._mean, self._sdev)

    def __repr__(self):
        return self.__str__()

    def __format__(self, spec):
        return format(str(self), spec)

    def __float__(self):
        return self._mean


class RAvg(GVar):
    """Running average of scalar Monte Carlo estimates (pyx:2276-2451):
    inverse-variance weighted when ``weighted`` (vegas' ``adapt=True``),
    plain average otherwise."""

    def __init__(self, weighted=True, itn_results=None, sum_neval=0):
        super().__init__(0.0, 0.0)
        self.weighted = weighted
        self._wlist = []
        self._mlist = []
        self._msum = 0.0
        self._varsum = 0.0
        self._n = 0
        self.itn_results = []
        if itn_results is not None:
            for r in itn_results:
                self.add(r)
        self.sum_neval = sum_neval

    def extend(self, ravg):
        """Merge results from ``ravg`` after those in ``self`` (pyx:2314-2318)."""
        for r in ravg.itn_results:
            self.add(r)
        self.sum_neval += ravg.sum_neval

    def __reduce__(self):
        return (_rebuild_ravg, (self.weighted, [(g.mean, g.sdev) for g in self.itn_results],
                                self.sum_neval))

    def add(self, g):
        """Add estimate ``g`` (pyx:2392-2410).  The variance re-read from g is
        sdev**2, as the reference re-reads it from its gvar."""
        self.add_all([g])

    def add_all(self, gs):
        """Add estimates in order; the same result as add() on each in turn
        (the weighted sums are formed once, over the whole list)."""
        for g in gs:
            if not isinstance(g, GVar):
                g = GVar(*g)
            self.itn_results.append(g)
            self._mlist.append(g.mean)
            if self.weighted:
                self._wlist.append(1 / (g.var if g.var > TINY else TINY))
            else:
                self._msum += g.mean
                self._varsum += g.var
                self._n += 1
        if not self.itn_results:
            return
        if self.weighted:
            var = 1. / np.sum(self._wlist)
            sdev = np.sqrt(var)
            mean = np.sum([w * m for w, m in zip(self._wlist, self._mlist)]) * var
        else:
            mean = self._msum / self._n
            sdev = np.sqrt(self._varsum / self._n ** 2)
        self._mean = float(mean)
        self._sdev = float(sdev)

    @property
    def chi2(self):
        """pyx:2344-2357"""
        if len(self.itn_results) <= 1:
            return 0.0
        if self.weighted:
            ans = 0.0
            for m, w in zip(self._mlist, self._wlist):
                ans += (self._mean - m) ** 2 * w
            return ans
        return np.sum([(m - self._mean) ** 2 for m in self._mlist]) / (self._varsum / self._n)

    @property
    def dof(self):
        return len(self.itn_results) - 1

    @property
    def nitn(self):
        return len(self.itn_results)

    @property
    def Q(self):
        """pyx:2372-2377"""
        return gammaQ(self.dof / 2., self.chi2 / 2.) if self.dof > 0 and self.chi2 >= 0 \
            else float('nan')

    @property
    def avg_neval(self):
        return self.sum_neval / self.nitn if self.nitn > 0 else 0

    def converged(self, rtol, atol):
        """pyx:2389-2390"""
        return self.sdev < atol + rtol * abs(self.mean)

    def summary(self, extended=False, weighted=None):
        """Iteration-by-iteration table (pyx:2412-2451)."""
        if weighted is None:
            weighted = self.weighted
        acc = RAvg(weighted=weighted)
        linedata = []
        for i, res in enumerate(self.itn_results):
            acc.add(res)
            itn = '%3d' % (i + 1)
            integral = '%-15s' % res
            wgtavg = '%-15s' % acc
            chi2dof = '%8.2f' % (acc.chi2 / acc.dof if i != 0 else 0.0)
            Q = '%8.2f' % (acc.Q if i != 0 else 1.0)
            linedata.append((itn, integral, wgtavg, chi2dof, Q))
        nchar = 5 * [0]
        for data in linedata:
            for i, d in enumerate(data):
                nchar[i] = max(nchar[i], len(d))
        fmt = '%%%ds   %%-%ds %%-%ds %%%ds %%%ds\n' % tuple(nchar)
        ans = fmt % ('itn', 'integral', 'wgt average' if weighted else 'average', 'chi2/dof', 'Q')
        ans += len(ans[:-1]) * '-' + '\n'
        for data in linedata:
            ans += fmt % data
        return ans


def _rebuild_ravg(weighted, results, sum_neval):
    return RAvg(weighted=weighted, itn_results=[GVar(m, s) for m, s in results],
                sum_neval=sum_neval)


# ------------------------------------------------------------------ arrays --
def _inv_sqrt(matrix, svdcut):
    """Rows w_i = v_i / sqrt(lambda_i) of the eigen-decomposition of the
    symmetric ``matrix`` (so that sum_i w_i w_i^T is its inverse on the kept
    modes).  svdcut < 0 drops the modes with lambda < |svdcut| max(lambda),
    as gvar.SVD(matrix, svdcut, rescale=False).decomp(-1) does for the
    reference's RAvgArray._w (pyx:2715-2721); svdcut <= 0 here."""
    matrix = np.asarray(matrix, np.float64)
    vals, vecs = np.linalg.eigh(matrix)
    top = vals.max() if vals.size else 0.0
    keep = (vals > abs(svdcut) * top) & (vals > 0)
    return (vecs[:, keep] / np.sqrt(vals[keep])).T


class GVarArray:
    """Estimates of several integrals: ``mean``, ``cov`` (covariance matrix)
    and ``sdev`` arrays of shape ``shape``; ``self[i]`` is a GVar."""

    def __init__(self, mean, cov, shape=None):
        mean = np.asarray(mean, np.float64).reshape(-1)
        self._m = mean
        self._cov = np.asarray(cov, np.float64).reshape(len(mean), len(mean))
        self.shape = tuple(shape) if shape is not None else (len(mean),)

    @property
    def mean(self):
        return self._m.reshape(self.shape)

    @property
    def cov(self):
        return self._cov

    @property
    def sdev(self):
        return np.sqrt(np.maximum(np.diag(self._cov), 0.0)).reshape(self.shape)

    @property
    def var(self):
        return self.sdev ** 2

    @property
    def size(self):
        return len(self._m)

    def flat_gvars(self):
        sd = np.sqrt(np.maximum(np.diag(self._cov), 0.0))
        return [GVar(m, s) for m, s in zip(self._m, sd)]

    @property
    def flat(self):
        """The estimates in C order, as a 1-d object array of GVars."""
        g = np.empty(self.size, dtype=object)
        g[:] = self.flat_gvars()
        return g

    def __getitem__(self, idx):
        g = np.empty(self.size, dtype=object)
        g[:] = self.flat_gvars()
        return g.reshape(self.shape)[idx]

    def __len__(self):
        return self.shape[0]

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    def __str__(self):
        g = np.empty(self.size, dtype=object)
        g[:] = [str(x) for x in self.flat_gvars()]
        return str(g.reshape(self.shape).tolist()).replace("'", "")

    __repr__ = __str__


class RAvgArray(GVarArray):
    """Running average of array-valued estimates (pyx:2453-2831).  weighted:
    the covariance-weighted average, each iteration's estimate rescaled by
    |mean| (or sdev where that is larger) of the first one and decomposed
    with the svdcut -EPSILON*len*1e4 of the reference; otherwise the plain
    average with covariance sum(cov)/n^2.  The per-iteration estimates
    ``itn_results`` are GVarArrays."""

    def __init__(self, shape, weighted=True, itn_results=None, sum_neval=0, rescale=True):
        n = int(np.prod(shape)) if len(shape) else 1
        super().__init__(np.zeros(n), np.zeros((n, n)), shape)
        self.weighted = weighted
        self.rescale = rescale
        self._mlist, self._wlist = [], []
        self._msum, self._covsum, self._n = 0.0, 0.0, 0
        self._invw = None
        self.itn_results = []
        self.sum_neval = sum_neval
        for r in itn_results or []:
            self.add(r)

    def _w(self, matrix):
        return _inv_sqrt(matrix, -EPSILON * len(matrix) * 1e4)

    def add(self, g):
        """Add the estimate ``g`` (a GVarArray, or (mean, cov)) (pyx:2775-2811)."""
        if not isinstance(g, GVarArray):
            g = GVarArray(g[0], g[1], self.shape)
        self.itn_results.append(g)
        gm = g._m.copy()
        gc = g._cov.copy()
        if self.weighted:
            if not hasattr(self, '_rescale'):
                if self.rescale is not None:
                    rs = np.fabs(gm)
                    gsd = np.sqrt(np.maximum(np.diag(gc), 0.0))
                    idx = gsd > rs
                    rs[idx] = gsd[idx]
                    rs[rs <= 0] = 1.
                    self._rescale = rs
                else:
                    self._rescale = np.ones(len(gm))
            rs = self._rescale
            gm = gm / rs
            gc = gc / np.outer(rs, rs)
            d = np.diag(gc).copy()
            d[d <= 0] = TINY
            gc[np.diag_indices_from(gc)] = d
            self._mlist.append(gm)
            self._wlist.append(self._w(gc))
            invcov = np.sum([w.T.dot(w) for w in self._wlist], axis=0)
            invw = self._w(invcov)
            cov = invw.T.dot(invw)
            mean = np.zeros(len(gm))
            for m, w in zip(self._mlist, self._wlist):
                for wj in w:
                    wj_m = wj.dot(m)
                    for invwi in invw:
                        mean += invwi * invwi.dot(wj) * wj_m
            self._m = mean * rs
            self._cov = cov * np.outer(rs, rs)
        else:
            self._mlist.append(gm)
            self._msum = self._msum + gm
            self._covsum = self._covsum + gc
            self._invw = None
            self._n += 1
            self._m = self._msum / self._n
            self._cov = self._covsum / self._n ** 2

    def extend(self, ravg):
        for r in ravg.itn_results:
            self.add(r)
        self.sum_neval += ravg.sum_neval

    @property
    def chi2(self):
        """pyx:2727-2748"""
        if len(self.itn_results) <= 1:
            return 0.0
        if self.weighted:
            wavg = self._m / self._rescale
            ans = 0.0
            for w, m in zip(self._wlist, self._mlist):
                for wi in w:
                    ans += wi.dot(m - wavg) ** 2
            return float(ans)
        if self._invw is None:
            self._invw = self._w(self._covsum / self._n)
        ans = 0.0
        for m in self._mlist:
            delta = self._m - m
            for invwi in self._invw:
                ans += invwi.dot(delta) ** 2
        return float(ans)

    @property
    def dof(self):
        """pyx:2751-2758"""
        if len(self.itn_results) <= 1:
            return 0
        if not self.weighted:
            if self._invw is None:
                self._invw = self._w(self._covsum / self._n)
            return (len(self.itn_results) - 1) * len(self._invw)
        return int(np.sum([len(w) for w in self._wlist]) - self.size)

    @property
    def nitn(self):
        return len(self.itn_results)

    @property
    def Q(self):
        if self.dof <= 0 or self.chi2 < 0:
            return float('nan')
        return gammaQ(self.dof / 2., self.chi2 / 2.)

    @property
    def avg_neval(self):
        return self.sum_neval / self.nitn if self.nitn > 0 else 0

    def converged(self, rtol, atol):
        """pyx:2723-2726"""
        return bool(np.all(self.sdev < atol + rtol * np.abs(self.mean)))

    def summary(self, extended=False, weighted=None, rescale=None):
        """Iteration-by-iteration table of the first component and the
        chi2 of the whole array (pyx:2833-2875); with ``extended`` a table of
        every component's average follows."""
        if weighted is None:
            weighted = self.weighted
        if rescale is None:
            rescale = self.rescale
        acc = RAvgArray(self.shape, weighted=weighted, rescale=rescale)
        linedata = []
        for i, res in enumerate(self.itn_results):
            acc.add(res)
            itn = '%3d' % (i + 1)
            integral = '%-15s' % res.flat_gvars()[0]
            wgtavg = '%-15s' % acc.flat_gvars()[0]
            chi2dof = '%8.2f' % (acc.chi2 / acc.dof if i != 0 else 0.0)
            Q = '%8.2f' % (acc.Q if i != 0 else 1.0)
            linedata.append((itn, integral, wgtavg, chi2dof, Q))
        nchar = 5 * [0]
        for data in linedata:
            for i, d in enumerate(data):
                nchar[i] = max(nchar[i], len(d))
        fmt = '%%%ds   %%-%ds %%-%ds %%%ds %%%ds\n' % tuple(nchar)
        ans = fmt % ('itn', 'integral', 'wgt average' if weighted else 'average', 'chi2/dof', 'Q')
        ans += len(ans[:-1]) * '-' + '\n'
        for data in linedata:
            ans += fmt % data
        if extended and self.size > 1:
            ans += '\n' + '\n'.join('%8s  %s' % (str(idx), g)
                                    for idx, g in zip(np.ndindex(self.shape), self.flat_gvars())) + '\n'
        return ans

    def __reduce__(self):
        return (_rebuild_ravg_array, (self.shape, self.weighted, self.rescale,
                                      [(r._m, r._cov) for r in self.itn_results], self.sum_neval))


def _rebuild_ravg_array(shape, weighted, rescale, results, sum_neval):
    r = RAvgArray(shape, weighted=weighted, rescale=rescale)
    for m, c in results:
        r.add(GVarArray(m, c, shape))
    r.sum_neval = sum_neval
    return r


class RAvgDict(dict):
    """Running average of dictionary-valued estimates: one RAvgArray over the
    flattened values (keys in the integrand's order), viewed per key."""

    def __init__(self, layout, weighted=True):
        super().__init__()
        self.layout = layout  # [(key, shape, start, stop)]
        n = layout[-1][3] if layout else 0
        self.rarray = RAvgArray((n,), weighted=weighted)
        self._refresh()

    def _refresh(self):
        for k, shape, a, b in self.layout:
            if shape == ():
                self[k] = GVar(self.rarray._m[a], math.sqrt(max(self.rarray._cov[a, a], 0.0)))
            else:
                self[k] = GVarArray(self.rarray._m[a:b], self.rarray._cov[a:b, a:b], shape)

    def add(self, g):
        self.rarray.add(g)
        self._refresh()

    def extend(self, ravg):
        """Merge results from ``ravg`` after those in ``self`` (pyx:2314-2318)."""
        self.rarray.extend(ravg.rarray)
        self._refresh()

    @property
    def flat(self):
        """Every key's estimates, keys in order, as a 1-d object array."""
        return self.rarray.flat

    @property
    def itn_results(self):
        return self.rarray.itn_results

    @property
    def sum_neval(self):
        return self.rarray.sum_neval

    @sum_neval.setter
    def sum_neval(self, v):
        self.rarray.sum_neval = v

    chi2 = property(lambda self: self.rarray.chi2)
    dof = property(lambda self: self.rarray.dof)
    Q = property(lambda self: self.rarray.Q)
    nitn = property(lambda self: self.rarray.nitn)

    def converged(self, rtol, atol):
        return self.rarray.converged(rtol, atol)

    def summary(self, extended=False, weighted=None):
        return self.rarray.summary(extended=extended, weighted=weighted)

    def __reduce__(self):
        return (_rebuild_ravg_dict, (self.layout, self.rarray))


def _rebuild_ravg_dict(layout, rarray):
    r = RAvgDict(layout, weighted=rarray.weighted)
    r.rarray = rarray
    r._refresh()
    return r


def ravg(reslist, weighted=None, rescale=None):
    """A running average of a list of vegas results (src/vegas/__init__.py:
    1220-1312): GVars give an RAvg, GVarArrays an RAvgArray of their shape,
    dictionaries of GVars / GVarArrays an RAvgDict (covariance block-diagonal
    across keys, as independent entries).  ``reslist`` may also be an RAvg,
    RAvgArray or RAvgDict, whose iteration results are averaged again.
    weighted defaults to the input's own setting (True for a plain list);
    rescale applies to array averages (True: by the last result)."""
    if isinstance(reslist, RAvgDict):
        if weighted is None:
            weighted = reslist.rarray.weighted
        out = RAvgDict(reslist.layout, weighted=weighted)
        out.rarray = RAvgArray(reslist.rarray.shape, weighted=weighted, itn_results=reslist.rarray.itn_results,
                               rescale=reslist.rarray.rescale if rescale is None else rescale)
        out._refresh()
        return out
    if isinstance(reslist, (RAvg, RAvgArray)):
        if weighted is None:
            weighted = reslist.weighted
        if rescale is None:
            rescale = getattr(reslist, 'rescale', True)
        reslist = reslist.itn_results
    try:
        if len(reslist) < 1:
            raise ValueError('reslist empty')
    except TypeError:
        raise ValueError('improper type for reslist')
    if weighted is None:
        weighted = True
    if rescale is None:
        rescale = True
    first = reslist[0]
    if hasattr(first, 'keys'):
        layout, a = [], 0
        for k, v in first.items():
            shape = () if isinstance(v, GVar) else tuple(np.shape(np.asarray(v.mean)))
            m = int(np.prod(shape)) if shape else 1
            layout.append((k, shape, a, a + m))
            a += m
        out = RAvgDict(layout, weighted=weighted)
        out.rarray = RAvgArray((a,), weighted=weighted, rescale=rescale)
        for r in reslist:
            mean = np.zeros(a)
            cov = np.zeros((a, a))
            for k, shape, lo, hi in layout:
                v = r[k]
                if isinstance(v, GVar):
                    mean[lo] = v.mean
                    cov[lo, lo] = v.sdev ** 2
                else:
                    mean[lo:hi] = np.asarray(v.mean, float).reshape(-1)
                    cov[lo:hi, lo:hi] = np.asarray(v.cov, float).reshape(hi - lo, hi - lo)
            out.rarray.add(GVarArray(mean, cov, (a,)))
        out._refresh()
        return out
    if isinstance(first, GVar) or np.shape(getattr(first, 'mean', first)) == ():
        return RAvg(weighted=weighted, itn_results=[r if isinstance(r, GVar) else GVar(*r) for r in reslist])
    if isinstance(first, GVarArray):
        return RAvgArray(first.shape, weighted=weighted, itn_results=list(reslist), rescale=rescale)
    raise ValueError('reslist[i] not GVar, GVarArray, or dictionary')

