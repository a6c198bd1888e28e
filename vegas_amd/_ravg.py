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


def _scalar_stats(means, variances, weighted):
    """Average, its standard deviation, chi^2 and degrees of freedom of the
    scalar estimates (means[k], variances[k]).

    weighted: inverse-variance weights, each variance floored at TINY (the
    reference's RAvg.add, pyx:2392-2410; chi^2 pyx:2344-2357);
    otherwise the plain average, whose variance is sum(var)/n^2 and whose
    chi^2 measures the spread against the mean variance sum(var)/n."""
    m = np.asarray(means, np.float64)
    v = np.asarray(variances, np.float64)
    n = m.size
    if n == 0:
        return 0.0, 0.0, 0.0, 0
    if weighted:
        w = 1.0 / np.maximum(v, TINY)
        wsum = w.sum()
        avg = float(np.dot(w, m) / wsum)
        sdev = float(math.sqrt(1.0 / wsum))
        chi2 = float(np.dot(w, (m - avg) ** 2)) if n > 1 else 0.0
    else:
        avg = float(m.sum() / n)
        vsum = float(v.sum())
        sdev = math.sqrt(vsum) / n
        with np.errstate(divide='ignore', invalid='ignore'):
            chi2 = float(np.sum((m - avg) ** 2) / np.float64(vsum / n)) if n > 1 else 0.0
    return avg, sdev, chi2, n - 1


def _q_value(dof, chi2):
    """p-value of chi2 with dof degrees of freedom (nan when undefined)."""
    if dof <= 0 or not chi2 >= 0:
        return float('nan')
    return gammaQ(dof / 2.0, chi2 / 2.0)


def _progress_table(rows, weighted):
    """The iteration table printed by the reference's summary() (pyx:2412-
    2451): a right-aligned iteration number, then the iteration's result, the
    running average (both left-aligned), chi2/dof and Q (right-aligned); the
    column widths are those of the widest entry of the rows below the header,
    and a rule as long as the header line separates them."""
    heads = ('itn', 'integral', 'wgt average' if weighted else 'average', 'chi2/dof', 'Q')
    widths = [max((len(r[c]) for r in rows), default=0) for c in range(5)]

    def line(cells):
        left = cells[0].rjust(widths[0])
        mid = [cells[1].ljust(widths[1]), cells[2].ljust(widths[2]),
               cells[3].rjust(widths[3]), cells[4].rjust(widths[4])]
        return left + '   ' + ' '.join(mid) + '\n'

    top = line(heads)
    return top + '-' * (len(top) - 1) + '\n' + ''.join(line(r) for r in rows)


def _progress_row(k, result, average, chi2, dof):
    """One row of the iteration table: k counts from 0; the first row shows
    chi2/dof 0 and Q 1 as the reference's does."""
    if k == 0:
        cd, q = 0.0, 1.0
    else:
        with np.errstate(divide='ignore', invalid='ignore'):
            cd = np.float64(chi2) / dof  # dof 0 (degenerate arrays): inf or nan
        q = _q_value(dof, chi2)
    return ('%3d' % (k + 1), '%-15s' % result, '%-15s' % average, '%8.2f' % cd, '%8.2f' % q)


class RAvg(GVar):
    """Running average of scalar Monte Carlo estimates (the reference's RAvg,
    pyx:2276-2451): inverse-variance weighted when ``weighted`` (vegas'
    ``adapt=True``), the plain average otherwise.

    The only state is the list of iteration results; the average, chi^2 and
    Q are derived from it by _scalar_stats whenever a result is added."""

    def __init__(self, weighted=True, itn_results=None, sum_neval=0):
        super().__init__(0.0, 0.0)
        self.weighted = weighted
        self.itn_results = []
        self._chi2 = 0.0
        if itn_results is not None:
            self.add_all(itn_results)
        self.sum_neval = sum_neval

    def extend(self, ravg):
        """Merge results from ``ravg`` after those in ``self`` (pyx:2314-2318)."""
        self.add_all(ravg.itn_results)
        self.sum_neval += ravg.sum_neval

    def __reduce__(self):
        return (_rebuild_ravg, (self.weighted, [(g.mean, g.sdev) for g in self.itn_results],
                                self.sum_neval))

    def add(self, g):
        """Add estimate ``g`` (a GVar or (mean, sdev)); its variance is sdev**2,
        as the reference re-reads it from its gvar (pyx:2392-2410)."""
        self.add_all([g])

    def add_all(self, gs):
        """Add estimates in order (the same result as add() on each)."""
        for g in gs:
            self.itn_results.append(g if isinstance(g, GVar) else GVar(*g))
        if self.itn_results:
            means = [g.mean for g in self.itn_results]
            variances = [g.var for g in self.itn_results]
            self._mean, self._sdev, self._chi2, _ = _scalar_stats(means, variances, self.weighted)

    @property
    def chi2(self):
        """chi^2 of the average (pyx:2344-2357)."""
        return self._chi2

    @property
    def dof(self):
        return len(self.itn_results) - 1

    @property
    def nitn(self):
        return len(self.itn_results)

    @property
    def Q(self):
        """p-value of chi2 (pyx:2372-2377)."""
        return _q_value(self.dof, self.chi2)

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
        means = [g.mean for g in self.itn_results]
        variances = [g.var for g in self.itn_results]
        rows = []
        for k, res in enumerate(self.itn_results):
            avg, sdev, chi2, dof = _scalar_stats(means[:k + 1], variances[:k + 1], weighted)
            rows.append(_progress_row(k, res, GVar(avg, sdev), chi2, dof))
        return _progress_table(rows, weighted)


def _rebuild_ravg(weighted, results, sum_neval):
    return RAvg(weighted=weighted, itn_results=[GVar(m, s) for m, s in results],
                sum_neval=sum_neval)


# ------------------------------------------------------------------ arrays --
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


def _pinv_psd(matrix, cut):
    """Pseudo-inverse of the symmetric positive semi-definite ``matrix``
    restricted to its eigenmodes with eigenvalue > cut * the largest one (and
    > 0), and the number of modes kept.  With cut = EPSILON * len * 1e4 this
    keeps the modes that gvar.SVD(matrix, svdcut=-cut).decomp(-1) keeps for
    the reference's RAvgArray._w (pyx:2715-2721)."""
    matrix = np.asarray(matrix, np.float64)
    if matrix.size == 0:
        return matrix.copy(), 0
    vals, vecs = np.linalg.eigh(matrix)
    keep = (vals > cut * vals.max()) & (vals > 0)
    v = vecs[:, keep]
    return (v / vals[keep]) @ v.T, int(keep.sum())


class RAvgArray(GVarArray):
    """Running average of array-valued estimates (pyx:2453-2831).

    weighted: the generalised least-squares average.  Each iteration's
    estimate is first divided by a fixed scale per component (|mean| of the
    first estimate, its sdev where that is larger, 1 where it is zero; or the
    array given as ``rescale``, or 1 with ``rescale=None``).  Its precision
    matrix P_k is the pseudo-inverse of its covariance (variances <= 0
    replaced by TINY) on the modes above the reference's eigenvalue cut
    EPSILON*len*1e4; the average is H^+ sum_k P_k m_k with H = sum_k P_k, its
    covariance H^+ (same cut), chi^2 = sum_k (m_k - avg)^T P_k (m_k - avg)
    and dof = sum_k rank(P_k) - len.
    Otherwise: the plain average, covariance sum(cov)/n^2, chi^2 against the
    mean covariance sum(cov)/n and dof = (n - 1) rank of that covariance.

    State: the iteration results, the scale, the running sums H and
    sum_k P_k m_k, and (m_k, P_k, rank_k) per iteration for chi^2."""

    def __init__(self, shape, weighted=True, itn_results=None, sum_neval=0, rescale=True):
        n = int(np.prod(shape)) if len(shape) else 1
        super().__init__(np.zeros(n), np.zeros((n, n)), shape)
        self.weighted = weighted
        self.rescale = rescale
        self._scale = None
        self._terms = []          # weighted: (scaled mean, precision, rank) per iteration
        self._prec_sum = np.zeros((n, n))
        self._info_sum = np.zeros(n)
        self._cov_sum = np.zeros((n, n))  # unweighted: sum of covariances
        self.itn_results = []
        self.sum_neval = sum_neval
        for r in itn_results or []:
            self.add(r)

    def _cut(self):
        return EPSILON * self.size * 1e4

    def _component_scale(self, g):
        if self.rescale is None or self.rescale is False:
            return np.ones(g.size)
        if self.rescale is True:
            scale = np.abs(g._m)
            sd = np.sqrt(np.maximum(np.diag(g._cov), 0.0))
            scale = np.where(sd > scale, sd, scale)
        else:
            scale = np.abs(np.asarray(self.rescale, np.float64).reshape(-1))
        return np.where(scale > 0, scale, 1.0)

    def add(self, g):
        """Add the estimate ``g`` (a GVarArray, or (mean, cov)) (pyx:2775-2811)."""
        if not isinstance(g, GVarArray):
            g = GVarArray(g[0], g[1], self.shape)
        self.itn_results.append(g)
        n = len(self.itn_results)
        if not self.weighted:
            self._cov_sum = self._cov_sum + g._cov
            self._m = np.mean([r._m for r in self.itn_results], axis=0)
            self._cov = self._cov_sum / n ** 2
            return
        if self._scale is None:
            self._scale = self._component_scale(g)
        s = self._scale
        mk = g._m / s
        ck = g._cov / np.outer(s, s)
        diag = np.diag(ck)
        ck = ck + np.diag(np.where(diag > 0, 0.0, TINY - diag))
        pk, rk = _pinv_psd(ck, self._cut())
        self._terms.append((mk, pk, rk))
        self._prec_sum = self._prec_sum + pk
        self._info_sum = self._info_sum + pk @ mk
        cov, _ = _pinv_psd(self._prec_sum, self._cut())
        self._m = (cov @ self._info_sum) * s
        self._cov = cov * np.outer(s, s)

    def extend(self, ravg):
        for r in ravg.itn_results:
            self.add(r)
        self.sum_neval += ravg.sum_neval

    def _mean_cov_pinv(self):
        pinv, rank = _pinv_psd(self._cov_sum / len(self.itn_results), self._cut())
        return pinv, rank

    @property
    def chi2(self):
        """pyx:2727-2748"""
        if len(self.itn_results) <= 1:
            return 0.0
        if self.weighted:
            avg = self._m / self._scale
            return float(sum(float((mk - avg) @ pk @ (mk - avg)) for mk, pk, _ in self._terms))
        pinv, _ = self._mean_cov_pinv()
        dev = np.array([r._m for r in self.itn_results]) - self._m
        return float(np.einsum('ki,ij,kj->', dev, pinv, dev))

    @property
    def dof(self):
        """pyx:2751-2758"""
        if len(self.itn_results) <= 1:
            return 0
        if self.weighted:
            return int(sum(rk for _, _, rk in self._terms) - self.size)
        return (len(self.itn_results) - 1) * self._mean_cov_pinv()[1]

    @property
    def nitn(self):
        return len(self.itn_results)

    @property
    def Q(self):
        return _q_value(self.dof, self.chi2)

    @property
    def avg_neval(self):
        return self.sum_neval / self.nitn if self.nitn > 0 else 0

    def converged(self, rtol, atol):
        """pyx:2723-2726"""
        return bool(np.all(self.sdev < atol + rtol * np.abs(self.mean)))

    def summary(self, extended=False, weighted=None, rescale=None):
        """Iteration-by-iteration table of the first component with the
        chi2 of the whole array (pyx:2833-2875); with ``extended`` a table of
        every component's average follows."""
        if weighted is None:
            weighted = self.weighted
        if rescale is None:
            rescale = self.rescale
        running = RAvgArray(self.shape, weighted=weighted, rescale=rescale)
        rows = []
        for k, res in enumerate(self.itn_results):
            running.add(res)
            rows.append(_progress_row(k, res.flat_gvars()[0], running.flat_gvars()[0],
                                      running.chi2, running.dof))
        text = _progress_table(rows, weighted)
        if extended and self.size > 1:
            text += '\n' + '\n'.join('%8s  %s' % (str(idx), g)
                                     for idx, g in zip(np.ndindex(self.shape), self.flat_gvars())) + '\n'
        return text

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

