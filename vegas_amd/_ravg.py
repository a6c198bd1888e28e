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
        if not isinstance(g, GVar):
            g = GVar(*g)
        self.itn_results.append(g)
        self._mlist.append(g.mean)
        if self.weighted:
            self._wlist.append(1 / (g.var if g.var > TINY else TINY))
            var = 1. / np.sum(self._wlist)
            sdev = np.sqrt(var)
            mean = np.sum([w * m for w, m in zip(self._wlist, self._mlist)]) * var
        else:
            self._msum += g.mean
            self._varsum += g.var
            self._n += 1
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
