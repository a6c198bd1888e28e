"""Expectation values over a probability density, and y-space
restratification, on top of the device :class:`Integrator`.

``PDFIntegrator`` follows src/vegas/__init__.py:373-1188: the parameters p of
a Gaussian (mean, covariance) are written p = mean + L chiv with L from the
covariance's eigen-decomposition, each chiv axis is mapped to
theta = arctan(chiv / scale) in [-arctan(limit / scale), arctan(limit / scale)]
with a vegas grid pre-trained on the Gaussian (``_make_map``), and vegas
integrates pdf(p) and f(p) pdf(p) over theta together; E[f] is their ratio,
with the error propagated from the two integrals' covariance.  The reference
takes gvar objects for the parameters; gvar is not a dependency here, so the
parameters are given as ``(mean, cov)`` (or anything with ``.mean`` and
``.cov``, such as an ``RAvgArray`` result), and results are plain
``GVar``/``GVarArray`` objects carrying the covariance.

``restratify`` follows __init__.py:1314-1590 for plain integrators: it
integrates f and its distribution over ndy slices of every y axis, weights
each axis by the variance of dI/dy there, and re-divides the hypercubes
among the axes by those weights.
"""
from __future__ import annotations

import math

import numpy as np

from ._ravg import GVar, GVarArray, RAvgDict
from .integrator import Integrator
from .adaptive_map import AdaptiveMap
from .batchintegrand import lbatchintegrand

__all__ = ['PDFIntegrator', 'PDFEV', 'PDFEVArray', 'PDFEVDict', 'restratify']


class _GaussPDF:
    """p = mean + L chiv for a Gaussian with covariance cov (the role of
    gvar.PDF in the reference): L = D V sqrt(lambda), with D the standard
    deviations and (lambda, V) the eigen-decomposition of the correlation
    matrix, eigenvalues below svdcut * max(lambda) raised to that floor;
    dp_dchiv = |det L|."""

    def __init__(self, mean, cov, svdcut=1e-15):
        mean = np.asarray(mean, float)
        self.shape = mean.shape
        self.meanflat = mean.reshape(-1).copy()
        n = self.meanflat.size
        cov = np.asarray(cov, float).reshape(n, n)
        d = np.sqrt(np.diag(cov))
        if np.any(d <= 0):
            raise ValueError('parameter variances must be positive')
        corr = cov / np.outer(d, d)
        lam, vec = np.linalg.eigh(0.5 * (corr + corr.T))
        lam = np.maximum(lam, abs(svdcut) * lam.max())
        self.L = d[:, None] * vec * np.sqrt(lam)[None, :]
        self.dp_dchiv = float(np.prod(d) * np.prod(np.sqrt(lam)))
        self.size = n
        self.cov = cov

    def pflat(self, chiv):
        """Flat parameters (n, size) for chiv (n, size)."""
        return self.meanflat[None, :] + chiv @ self.L.T


def _param(param):
    if hasattr(param, 'mean') and hasattr(param, 'cov'):
        return np.asarray(param.mean, float), np.asarray(param.cov, float)
    mean, cov = param
    return np.asarray(mean, float), np.asarray(cov, float)


def _call_user(fn, p):
    """fn at parameter points p (n, *shape) in fn's batch convention;
    returns an (n, ...) array or a dict of them."""
    kind = getattr(fn, 'fcntype', 'scalar')
    if kind == 'lbatch':
        out = fn(p)
        if hasattr(out, 'keys'):
            return {k: np.asarray(v, float) for k, v in out.items()}
        return np.asarray(out, float)
    if kind == 'rbatch':
        out = fn(np.moveaxis(p, 0, -1))
        if hasattr(out, 'keys'):
            return {k: np.moveaxis(np.asarray(v, float), -1, 0) for k, v in out.items()}
        return np.moveaxis(np.asarray(out, float), -1, 0)
    outs = [fn(pi) for pi in p]
    if hasattr(outs[0], 'keys'):
        return {k: np.array([o[k] for o in outs], float) for k in outs[0]}
    return np.array(outs, float)


def _flatten_f(fp, n):
    """(n, m) columns and a rebuild layout for an array or dict output."""
    if isinstance(fp, dict):
        cols, layout, a = [], [], 0
        for k, v in fp.items():
            shape = v.shape[1:]
            m = int(np.prod(shape)) if shape else 1
            cols.append(v.reshape(n, m))
            layout.append((k, shape, a, a + m))
            a += m
        return np.concatenate(cols, axis=1), ('dict', layout)
    shape = fp.shape[1:]
    return fp.reshape(n, -1), ('array', shape)


class PDFEV(GVar):
    """Expectation value of a scalar f(p) (__init__.py:81-176): a GVar with
    ``pdfnorm`` (the integral of the PDF) and ``results`` (the running
    average of the two integrals), whose attributes (summary, chi2, Q, ...)
    it forwards."""

    def __init__(self, mean, sdev, results=None, pdfnorm=None):
        super().__init__(mean, sdev)
        self.results = results
        self.pdfnorm = pdfnorm

    def __getattr__(self, k):
        if k in ('results', 'pdfnorm') or k.startswith('__'):
            raise AttributeError(k)
        return getattr(self.results, k)


class PDFEVArray(GVarArray):
    """Expectation values of an array-valued f(p) (__init__.py:178-272)."""

    def __init__(self, mean, cov, shape, results=None, pdfnorm=None):
        super().__init__(mean, cov, shape)
        self.results = results
        self.pdfnorm = pdfnorm

    def __getattr__(self, k):
        if k in ('results', 'pdfnorm') or k.startswith('_'):
            raise AttributeError(k)
        return getattr(self.results, k)


class PDFEVDict(dict):
    """Expectation values of a dictionary-valued f(p) (__init__.py:274-371):
    one GVar or GVarArray per key; ``cov`` is the covariance of all values,
    keys in order."""

    def __init__(self, items, cov, results=None, pdfnorm=None):
        super().__init__(items)
        self.cov = cov
        self.results = results
        self.pdfnorm = pdfnorm

    def __getattr__(self, k):
        if k in ('results', 'pdfnorm', 'cov') or k.startswith('_'):
            raise AttributeError(k)
        return getattr(self.results, k)


def _view(g, a, b, shape):
    """Entries a..b of the flat GVarArray g as a GVar (shape ()) or a
    GVarArray of the given shape with their covariance block."""
    if shape == ():
        return GVar(g._m[a], math.sqrt(max(g._cov[a, a], 0.0)))
    return GVarArray(g._m[a:b], g._cov[a:b, a:b], shape)


def _ratio(mean, cov, ipdf):
    """E = F / P and its covariance to first order, from the means and
    covariance of (P, F_1..F_m) with P at index ipdf."""
    m = len(mean)
    fi = [i for i in range(m) if i != ipdf]
    P = mean[ipdf]
    F = mean[fi]
    J = np.zeros((len(fi), m))
    J[np.arange(len(fi)), fi] = 1.0 / P
    J[:, ipdf] = -F / P ** 2
    return F / P, J @ cov @ J.T


def _hist_counts(v, bins):
    """One-hot bin membership (n, k, nbin + 2) of each column of v (n, k):
    index 0 below bins[j][0], nbin + 1 at or above bins[j][-1], as
    searchsorted(side='right') does (__init__.py:1036-1042)."""
    n, k = v.shape
    nb2 = bins.shape[1] + 1
    out = np.zeros((n, k, nb2))
    rows = np.arange(n)
    for j in range(k):
        out[rows, j, np.searchsorted(bins[j], v[:, j], side='right')] = 1.0
    return out


class PDFStatistics:
    """Statistics of one component of f(p) from its first four raw moments
    m1..m4 under the PDF (gvar.PDFStatistics, as used at
    __init__.py:999-1010): ``mean``, ``sdev``, ``skew`` and ``ex_kurt``
    (excess kurtosis), each a GVar whose error is propagated to first order
    from the covariance ``cov`` of the moments' Monte Carlo estimates."""

    def __init__(self, moments, cov=None, histogram=None):
        nm = len(moments)
        if nm not in (2, 4):
            raise ValueError('moments: [m1, m2] or [m1, m2, m3, m4]')
        cov = np.zeros((nm, nm)) if cov is None else np.asarray(cov, float).reshape(nm, nm)
        if nm == 2:  # mean and sdev only
            moments = list(moments) + [0.0, 0.0]
            c4 = np.zeros((4, 4))
            c4[:2, :2] = cov
            cov = c4
        m1, m2, m3, m4 = (float(v) for v in moments)
        var = m2 - m1 * m1
        c3 = m3 - 3 * m1 * m2 + 2 * m1 ** 3
        c4 = m4 - 4 * m1 * m3 + 6 * m1 * m1 * m2 - 3 * m1 ** 4
        dvar = np.array([-2 * m1, 1.0, 0.0, 0.0])
        sd = math.sqrt(var) if var > 0 else 0.0
        vals = [(m1, np.array([1.0, 0.0, 0.0, 0.0]))]
        if var > 0:
            dc3 = np.array([-3 * m2 + 6 * m1 * m1, -3 * m1, 1.0, 0.0])
            dc4 = np.array([-4 * m3 + 12 * m1 * m2 - 12 * m1 ** 3, 6 * m1 * m1, -4 * m1, 1.0])
            vals += [(sd, dvar / (2 * sd)),
                     (c3 / var ** 1.5, dc3 / var ** 1.5 - 1.5 * c3 / var ** 2.5 * dvar),
                     (c4 / var ** 2 - 3.0, dc4 / var ** 2 - 2 * c4 / var ** 3 * dvar)]
        else:
            vals += [(0.0, np.zeros(4)), (float('nan'), np.zeros(4)), (float('nan'), np.zeros(4))]
        g = [GVar(v, math.sqrt(max(float(J @ cov @ J), 0.0))) for v, J in vals]
        self.mean, self.sdev = g[:2]
        self.skew, self.ex_kurt = (g[2], g[3]) if nm == 4 else (None, None)
        self.moments = [m1, m2, m3, m4][:nm]
        # histogram: bin edges [nbin + 1] and the probability of each bin,
        # below the first edge first and above the last edge last
        self.histogram = None
        self.median = None
        if histogram is not None:
            bins, prob, perr = (np.asarray(a, float) for a in histogram)
            self.histogram = (bins, prob, perr)
            # median: the CDF interpolated linearly inside the bin where it
            # passes one half (parity unpinned: gvar's own rule is not here)
            cdf = np.cumsum(prob[:-1])
            j = int(np.searchsorted(cdf, 0.5 * np.sum(prob)))
            if 1 <= j <= len(bins) - 1:
                c0 = cdf[j - 1]
                frac = (0.5 * np.sum(prob) - c0) / prob[j] if prob[j] > 0 else 0.5
                self.median = float(bins[j - 1] + frac * (bins[j] - bins[j - 1]))

    def __str__(self):
        out = 'mean = %s   sdev = %s' % (self.mean, self.sdev)
        if self.skew is not None:
            out += '   skew = %s   ex_kurt = %s' % (self.skew, self.ex_kurt)
        if self.median is not None:
            out += '   median = %.5g' % self.median
        return out

    __repr__ = __str__


class PDFIntegrator(Integrator):
    """Expectation values with respect to a PDF (__init__.py:373-1188).

    ``param``: ``(mean, cov)`` of the Gaussian that shapes the integration
    variables (or an object with ``.mean`` and ``.cov``); ``pdf``: the
    (unnormalised) density, a function of p in any batch convention, or
    None for the Gaussian itself; ``adapt_to_pdf``: vegas adapts to the
    PDF (True) or to f(p) pdf(p); ``limit``, ``scale``, ``svdcut``: as in
    the reference.  Other keywords go to :class:`Integrator`."""

    def __init__(self, param, pdf=None, adapt_to_pdf=True, limit=100., scale=1., svdcut=1e-15, **kargs):
        if isinstance(param, PDFIntegrator):
            super().__init__(param)
            for k in ('param_pdf', 'pdf', 'adapt_to_pdf', 'limit', 'scale', 'svdcut', '_param'):
                setattr(self, k, getattr(param, k))
            if kargs:
                self.set(kargs)
            return
        self._param = _param(param)
        self.param_pdf = _GaussPDF(*self._param, svdcut=svdcut)
        self.limit = abs(limit)
        self.scale = abs(scale)
        self.svdcut = svdcut
        self.pdf = pdf
        self.adapt_to_pdf = adapt_to_pdf
        grid = self._make_map(self.limit / self.scale, self.scale)
        kargs = {k: v for k, v in kargs.items() if k != 'uses_jac'}
        super().__init__(AdaptiveMap(self.param_pdf.size * [list(grid)]), **kargs)

    def __reduce__(self):
        _, _, odict = Integrator.__reduce__(self)
        return (PDFIntegrator, (self._param, self.pdf, self.adapt_to_pdf, self.limit, self.scale, self.svdcut),
                odict)

    def set(self, ka={}, **kargs):
        kargs = dict(ka, **kargs) if kargs else dict(ka)
        if 'param' in kargs:
            raise ValueError("Can't reset param.")
        old = {}
        for k in ('pdf', 'adapt_to_pdf'):
            if k in kargs:
                old[k] = getattr(self, k, None)
                setattr(self, k, kargs.pop(k))
        kargs.pop('uses_jac', None)
        if kargs:
            old.update(super().set(kargs))
        return old

    @staticmethod
    def _make_map(limit, scale=1.0):
        """Grid nodes in theta = arctan(chiv / scale) adapted to the unit
        Gaussian in chiv = scale * tan(theta) (__init__.py:576-592: 2000
        points, 100 increments, 10 passes of alpha 1.5; the training weight
        is dchiv/dtheta / scale times exp(-chiv^2 / 2))."""
        ny = 2000
        y = np.random.default_rng(20240601).random((ny, 1))
        lim = math.atan(limit)
        m = AdaptiveMap([[-lim, lim]], ninc=100)
        for _ in range(10):
            theta, jac = m.map(y)
            theta = theta.cpu().numpy()[:, 0]
            jac = jac.cpu().numpy()
            t = np.tan(theta)
            chiv = scale * t
            fx = (t ** 2 + 1) * np.exp(-(chiv ** 2) / 2.)
            m.add_training_data(y, (jac * fx) ** 2)
            m.adapt(alpha=1.5)
        return np.array(m.extract_grid()[0])

    def _p_and_weight(self, theta):
        """Parameters p (n, *shape) and dp/dtheta * pdf(p) at theta (n, size)
        (__init__.py:594-630)."""
        t = np.tan(theta)
        chiv = self.scale * t
        dp_dtheta = np.prod(self.scale * (t ** 2 + 1.), axis=1) * self.param_pdf.dp_dchiv
        pflat = self.param_pdf.pflat(chiv)
        p = pflat.reshape((len(theta),) + self.param_pdf.shape)
        if self.pdf is None:
            pdfv = np.prod(np.exp(-(chiv ** 2) / 2.) / math.sqrt(2 * math.pi), axis=1) / self.param_pdf.dp_dchiv
        else:
            pdfv = np.asarray(_call_user(self.pdf, p), float).reshape(len(theta), -1)
            pdfv = np.prod(pdfv, axis=1)
        return p, dp_dtheta * pdfv

    def __call__(self, f=None, save=None, saveall=None, **kargs):
        """E[f(p)] over the PDF (__init__.py:632-718): a PDFEV, PDFEVArray or
        PDFEVDict; with f=None the integral of the PDF alone (a running
        average keyed 'pdf', as the reference's RAvgDict, with ``pdfnorm``)."""
        kargs.pop('uses_jac', None)
        if kargs:
            self.set(kargs)
        layout = {}
        pdf_first = self.adapt_to_pdf or f is None

        @lbatchintegrand
        def integrand(theta):
            p, w = self._p_and_weight(theta)
            n = len(theta)
            if f is None:
                return w[:, None]
            cols, layout['l'] = _flatten_f(_call_user(f, p), n)
            fw = cols * w[:, None]
            return np.concatenate([w[:, None], fw] if pdf_first else [fw, w[:, None]], axis=1)

        results = Integrator.__call__(self, integrand, save=save, saveall=saveall)
        ipdf = 0 if pdf_first else len(results.mean) - 1
        pdfnorm = GVar(results.mean[ipdf], math.sqrt(max(results.cov[ipdf, ipdf], 0.0)))
        if pdfnorm.mean == 0:
            raise RuntimeError('Integral of PDF vanishes; increase neval?')
        if f is None:
            # the reference returns its RAvgDict of the one integral 'pdf'
            ans = RAvgDict([('pdf', (), 0, 1)], weighted=results.weighted)
            ans.rarray = results
            ans._refresh()
            ans.pdfnorm = pdfnorm
            return ans
        mean, cov = _ratio(np.asarray(results.mean, float).reshape(-1), results.cov, ipdf)
        kind, lay = layout['l']
        if kind == 'array':
            if lay == ():
                return PDFEV(mean[0], math.sqrt(max(cov[0, 0], 0.0)), results, pdfnorm)
            return PDFEVArray(mean, cov, lay, results, pdfnorm)
        items = {}
        for k, shape, a, b in lay:
            if shape == ():
                items[k] = GVar(mean[a], math.sqrt(max(cov[a, a], 0.0)))
            else:
                items[k] = GVarArray(mean[a:b], cov[a:b, a:b], shape)
        return PDFEVDict(items, cov, results, pdfnorm)

    def stats(self, f=None, moments=False, histograms=False, **kargs):
        """Mean and covariance of the distribution of f(p) under the PDF
        (__init__.py:729-946): E[f] and E[f f^T] - E[f] E[f]^T from one
        integration, plus the integration's own covariance of E[f]
        (__init__.py:965-968).  f=None means f(p) = p.  The grid is not
        adapted unless ``adapt`` is given (__init__.py:871-874).  The result
        (PDFEV, PDFEVArray or PDFEVDict, with ``results`` and ``pdfnorm``)
        carries, as the reference's does (__init__.py:948-1017):
        ``stats``, a :class:`PDFStatistics` per component of f from E[f] and
        E[f^2] (and E[f^3], E[f^4] with ``moments=True``, and the probability
        of each histogram bin with ``histograms``), all from the same
        integration; ``vegas_mean`` (E[f] with its integration error),
        ``vegas_cov`` (the covariance of f, with the error propagated from
        the integrals of f and f f^T) and ``vegas_sdev``.
        ``histograms`` is True or a dict with ``nbin`` (12), ``binwidth``
        (0.5) and ``loc`` = (mean, sdev) arrays of the histogram centres and
        widths; without ``loc`` they come from a one-iteration integration
        of f and f^2 (__init__.py:889-926)."""
        old_adapt = None
        if 'adapt' not in kargs:
            old_adapt = self.adapt
            kargs['adapt'] = False
        shape_box = {}

        def fvals(p):
            # f's values as (n, k) columns; a dictionary's keys in order
            fp = p if f is None else _call_user(f, p)
            n = len(p)
            if isinstance(fp, dict):
                v, shape_box['lay'] = _flatten_f(fp, n)
                return v
            fp = np.asarray(fp, float)
            shape_box['lay'] = ('array', fp.shape[1:])
            return fp.reshape(n, -1)

        hist = None
        if histograms is not False:
            hist = {} if histograms is True else dict(histograms)
            nbin = int(hist.get('nbin', 12))
            binwidth = float(hist.get('binwidth', 0.5))
            loc = hist.get('loc', None)
            if loc is not None:
                hmean = np.asarray(loc[0], float).reshape(-1)
                hsdev = np.asarray(loc[1], float).reshape(-1)
            else:
                @lbatchintegrand
                def ff2(p):
                    v = fvals(p)
                    return np.concatenate([v, v * v], axis=1)
                old_nitn = self.nitn
                r = self(ff2, nitn=1, **{k: v for k, v in kargs.items() if k != 'nitn'})
                self.set(nitn=old_nitn)
                m2 = np.asarray(r.mean, float).reshape(-1)
                kk = len(m2) // 2
                hmean = m2[:kk]
                hsdev = np.fabs(m2[kk:] - hmean * hmean) ** 0.5
            half = nbin / 2 * binwidth
            hist['bins'] = np.array([hmean[i] + np.linspace(-half * hsdev[i], half * hsdev[i], nbin + 1)
                                     for i in range(len(hmean))])
            hist['nbin'] = nbin

        @lbatchintegrand
        def fff(p):
            v = fvals(p)
            il = np.tril_indices(v.shape[1])
            cols = [v, (v[:, :, None] * v[:, None, :])[:, il[0], il[1]]]
            if moments:
                cols += [v ** 3, v ** 4]
            if hist is not None:
                cols.append(_hist_counts(v, hist['bins']).reshape(v.shape[0], -1))
            return np.concatenate(cols, axis=1)

        try:
            ev = self(fff, **kargs)
        finally:
            if old_adapt is not None:
                self.set(adapt=old_adapt)
        m = np.asarray(ev.mean, float).reshape(-1)
        C = np.asarray(ev.cov, float).reshape(len(m), len(m))
        kind, lay = shape_box['lay']
        shape = lay if kind == 'array' else None
        k = (int(np.prod(shape)) if shape else 1) if kind == 'array' else lay[-1][3]
        mu = m[:k]
        second = np.zeros((k, k))
        il = np.tril_indices(k)
        second[il] = m[k:k + len(il[0])]
        second = second + np.tril(second, -1).T
        cov = second - np.outer(mu, mu) + C[:k, :k]
        # vegas_cov: the covariance of f under the PDF as estimated, with its
        # first-order error from the integration covariance of (E[f],
        # E[f f^T]); vegas_sdev its diagonal's square root
        nt = len(il[0])
        tri = np.zeros((k, k), int)
        tri[il] = np.arange(nt)
        tri = np.maximum(tri, tri.T)
        jac = np.zeros((k * k, k + nt))
        for i in range(k):
            for j in range(k):
                r = i * k + j
                jac[r, k + tri[i, j]] += 1.0
                jac[r, i] -= mu[j]
                jac[r, j] -= mu[i]
        cov_err = jac @ C[:k + nt, :k + nt] @ jac.T
        dg = np.arange(k) * (k + 1)
        sdev_val = np.sqrt(np.fabs(np.diag(cov)))
        with np.errstate(divide='ignore', invalid='ignore'):
            jsd = np.where(sdev_val > 0, np.sign(np.diag(cov)) / (2 * sdev_val), 0.0)
        sdev_err = jsd[:, None] * cov_err[np.ix_(dg, dg)] * jsd[None, :]
        vmean = GVarArray(mu, C[:k, :k], (k,))
        vcov = GVarArray(cov.reshape(-1), cov_err, (k, k))
        vsdev = GVarArray(sdev_val, sdev_err, (k,))
        # PDFStatistics per component: always from the first two moments,
        # plus the third and fourth with moments=True, and the histogram
        n2 = k + nt
        n4 = n2 + (2 * k if moments else 0)
        st = []
        for i in range(k):
            idx = [i, k + i * (i + 1) // 2 + i] + ([n2 + i, n2 + k + i] if moments else [])
            h = None
            if hist is not None:
                nb2 = hist['nbin'] + 2
                hi = np.arange(n4 + i * nb2, n4 + (i + 1) * nb2)
                h = (hist['bins'][i], m[hi], np.sqrt(np.maximum(np.diag(C)[hi], 0.0)))
            st.append(PDFStatistics(m[idx], C[np.ix_(idx, idx)], histogram=h))
        results, pdfnorm = ev.results, ev.pdfnorm
        if kind == 'dict':
            items, vm, vs, vc = {}, {}, {}, {}
            for key, shp, a, b in lay:
                items[key] = (GVar(mu[a], math.sqrt(max(cov[a, a], 0.0))) if shp == ()
                              else GVarArray(mu[a:b], cov[a:b, a:b], shp))
                vm[key] = _view(vmean, a, b, shp)
                vs[key] = _view(vsdev, a, b, shp)
                for key2, shp2, a2, b2 in lay:
                    rows = (np.arange(a, b)[:, None] * k + np.arange(a2, b2)[None, :]).reshape(-1)
                    blk = tuple(shp) + tuple(shp2)  # gvar.evalcov's block shape
                    vc[key, key2] = (GVar(cov[a, a2], math.sqrt(max(cov_err[rows[0], rows[0]], 0.0)))
                                     if blk == () else
                                     GVarArray(cov[a:b, a2:b2].reshape(-1), cov_err[np.ix_(rows, rows)], blk))
            ans = PDFEVDict(items, cov, results, pdfnorm)
            ans.stats = {key: (st[a] if shp == () else np.array(st[a:b], dtype=object).reshape(shp))
                         for key, shp, a, b in lay}
            ans.vegas_mean, ans.vegas_cov, ans.vegas_sdev = vm, vc, vs
        elif shape == ():
            ans = PDFEV(mu[0], math.sqrt(max(cov[0, 0], 0.0)), results, pdfnorm)
            ans.stats = st[0]
            ans.vegas_mean = GVar(mu[0], math.sqrt(max(C[0, 0], 0.0)))
            ans.vegas_cov = GVar(cov[0, 0], math.sqrt(max(cov_err[0, 0], 0.0)))
            ans.vegas_sdev = GVar(sdev_val[0], math.sqrt(max(sdev_err[0, 0], 0.0)))
        else:
            ans = PDFEVArray(mu, cov, shape, results, pdfnorm)
            ans.stats = np.array(st, dtype=object).reshape(shape)
            ans.vegas_mean = GVarArray(mu, C[:k, :k], shape)
            ans.vegas_cov = GVarArray(cov.reshape(-1), cov_err, shape + shape)
            ans.vegas_sdev = GVarArray(sdev_val, sdev_err, shape)
        return ans

    def sample(self, nbatch, mode='rbatch'):
        """Weighted samples of p from the PDF (__init__.py:1049-1188): at
        least nbatch samples (whole iterations), weights summing to one;
        the batch index last (``'rbatch'``) or first (``'lbatch'``)."""
        neval = self.last_neval if self.last_neval > 0 else self.neval
        nit = 1 if nbatch is None else int(nbatch) // neval
        if nbatch is not None and nit * neval < nbatch:
            nit += 1
        samples, wgts = [], []
        for _ in range(nit):
            for theta, wgt in self.random_batch():
                p, w = self._p_and_weight(theta)
                wgts.append(wgt * w)
                samples.append(p)
        samples = np.concatenate(samples, axis=0)
        wgts = np.concatenate(wgts)
        wgts /= np.sum(wgts)
        if mode == 'rbatch':
            samples = np.moveaxis(samples, 0, -1)
        return wgts, samples


def _restratified(weight, old_nstrat, gamma=1., below_avg_nstrat=None):
    """New strata per axis from the axis weights (the rule of
    __init__.py:1350-1366).

    The target of axis mu is old_mu * ((w_mu / gm(w)) * gm(old) / old_mu)
    ** gamma (gm: geometric mean, prod ** (1/dim)).  Axes are then fixed to
    integers in order of increasing target: truncation (at least 1), or
    ``below_avg_nstrat`` for an axis of below-average weight.  The ratio
    target / assigned of each fixed axis is spread over the targets of the
    axes still open, as the factor ratio ** (1 / number still open) on each,
    so the product of the strata follows the targets' product.  The
    arithmetic is done in the reference's order: where the product makes the
    last target an integer in exact arithmetic, its truncation depends on
    the rounding."""
    w = np.asarray(weight, float)
    old = np.asarray(old_nstrat)
    dim = len(old)
    gm_w = np.prod(w) ** (1 / dim)
    gm_old = np.prod(old) ** (1 / dim)
    target = old * ((w / gm_w) * gm_old / old) ** gamma
    below = w < np.average(w) if below_avg_nstrat else np.zeros(dim, bool)
    out = np.array(old_nstrat)
    order = np.argsort(target)
    for done, mu in enumerate(order, start=1):
        want = target[mu]
        out[mu] = below_avg_nstrat if below[mu] else (want if want > 1 else 1)
        still_open = order[done:]
        if len(still_open):
            target[still_open] *= (want / out[mu]) ** (1 / len(still_open))
    return out


def _report_strata(title, summary, nstrat, tail=''):
    """restratify(verbose=True): a section with the training summary and the
    strata (__init__.py:1340-1346, 1382-1388 print the same fields)."""
    nstrat = np.asarray(nstrat)
    lines = [title, summary[:-1],
             'nstrat = ' + np.array2string(nstrat, max_line_width=60, prefix=9 * ' '),
             'nhcube = %d%s' % (int(np.prod(nstrat)), tail)]
    print('\n'.join(lines))


def restratify(integ, f=None, nitn=1, ndy=5, below_avg_nstrat=None, verbose=False, gamma=1., **vargs):
    """A copy of ``integ`` whose strata per axis follow how much the
    integral of f varies along each y axis (__init__.py:1314-1590): weight
    w_mu = ndy * sum_i (dI_mu,i - I / ndy)^2 over ndy equal slices of y_mu;
    nstrat_mu scaled by ((w_mu / geometric mean of w) * geometric mean of
    nstrat / nstrat_mu) ** gamma, smallest first, the total kept by
    rescaling the remaining axes; axes of below-average weight get
    ``below_avg_nstrat`` strata when that is set.  The copy is then trained
    on f for nitn iterations.  For a :class:`PDFIntegrator` the integrand
    is the PDF itself, in the integrator's own variables, and the copy is a
    PDFIntegrator trained on the PDF (__init__.py:1369-1374, 1436-1456)."""
    is_pdf = isinstance(integ, PDFIntegrator)
    if is_pdf:
        # a copy through the pickle state, without serialising (the PDF
        # may be a local function): settings, sigf, grid and stream position
        cls, args, odict = integ.__reduce__()
        new = cls(*args)
        new.__setstate__(odict)
    else:
        if f is None:
            raise ValueError('restratify needs the integrand f')
        new = Integrator(integ)
    if vargs:
        new.set(vargs)
    dim = new.dim
    yst = np.linspace(0, 1, ndy + 1)
    old_nstrat = np.array(new.nstrat)

    @lbatchintegrand
    def I_dI(x):
        n = x.shape[0]
        y = np.zeros(x.shape, float)
        jac = np.zeros(n, float)
        new.map.invmap(x, y, jac)
        if is_pdf:
            I = new._p_and_weight(np.asarray(x, float))[1]
        else:
            I = np.asarray(_call_user(f, x), float).reshape(n, -1)[:, 0]
        dI = np.zeros((n, dim, ndy), float)
        for mu in range(dim):
            for i in range(ndy):
                idx = (yst[i] <= y[:, mu]) & (y[:, mu] <= yst[i + 1])
                dI[idx, mu, i] += I[idx]
        return dict(I=I, dI=dI)

    save = new.set(correlate_integrals=False, nitn=nitn, alpha=0, adapt=True)
    # in the integrator's own variables (for a PDFIntegrator: theta)
    result = Integrator.__call__(new, I_dI)
    new.set(save)
    if verbose:
        _report_strata('\n' + 20 * '=' + ' restratify\nBEFORE:', result.summary(), old_nstrat, '\n')
    # first-order gvar arithmetic takes means through: the weights are
    # functions of the means
    Im = float(np.asarray(result['I'].mean))
    dIm = np.asarray(result['dI'].mean).reshape(dim, ndy)
    weight = np.array([np.sum((dIm[mu] - Im / ndy) ** 2) * ndy for mu in range(dim)])
    if verbose:
        print('WEIGHTS:', weight)
    new_nstrat = _restratified(weight, old_nstrat, gamma, below_avg_nstrat)
    new.set(nstrat=new_nstrat, neval=new.neval)
    save = new.set(nitn=nitn, adapt=True)
    training = new() if is_pdf else new(f)
    new.set(save)
    new.restratify_weight = weight
    if verbose:
        _report_strata('\nAFTER:', training.summary(), new.nstrat, '\n' + 20 * '=')
    return new
