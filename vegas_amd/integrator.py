"""Integrator -- the reference's ``vegas.Integrator`` (src/vegas/_vegas.pyx:
834-2230) with every iteration executed on the GPU.

Two execution paths behind the same ``__call__``:

* device integrands (:mod:`vegas_amd.integrands`): one ``vgx_integrate`` call
  enqueues all ``nitn`` iterations -- allocation, sampling, map, integrand,
  per-hcube sums, training histogram, sigf update, adapt and the running
  average -- with no host synchronisation until the results are read;
* Python batch integrands (the reference's ``@lbatchintegrand`` /
  ``@rbatchintegrand`` convention, numpy or torch): the iteration is staged
  in batches cut at hcube boundaries exactly as ``_random_batch`` cuts them
  (pyx:1692-1710); the device samples each batch, the integrand is called on
  it, and the device accumulates it.

The random numbers come from Philox4x32-10 keyed by ``seed``, with counter
(point index within the iteration, iteration, axis block): the sample does
not depend on batching or launch configuration.

Options beyond the reference's: ``seed``; ``precision`` ('f32': 24-bit
uniforms and the FP32 integrand functors, the default for device
integrands; 'f64'); ``exact`` (FP64 parity mode: y by IEEE division, the
reference's sequential adapt, the training histogram summed in point order,
bit for bit the reference's for identical y); ``deterministic`` (default
True: the training histogram is summed in a fixed order, so a run repeats
bit for bit for a given seed on a given GPU; False adds points in the order
the device happens to take them, about 14% faster on the fused sampler for a
cheap integrand, 2% for an expensive one).
"""
from __future__ import annotations

import ctypes as C
import math
import pickle
import warnings

import numpy as np

from . import _geometry, _lib
from ._ravg import RAvg, RAvgArray, RAvgDict, GVar, GVarArray
from .adaptive_map import AdaptiveMap
from .integrands import DeviceIntegrand

HUGE = 10 ** (308 - 50)


def _region_layout(region):
    """Structure of the integration variables (the reference's xsample,
    pyx:1209-1253): None for a plain list of limits or node lists; for a
    dictionary, [(key, shape, start, stop)] over the flattened axes in key
    order; for an array of limits of more than one dimension, [(None,
    shape, 0, n)]."""
    if isinstance(region, (AdaptiveMap, Integrator)):
        return getattr(region, '_xlayout', None)
    if hasattr(region, 'keys'):
        layout, a = [], 0
        for k in region:
            shape = tuple(np.shape(np.asarray(region[k], float)))[:-1]
            n = int(np.prod(shape)) if shape else 1
            layout.append((k, shape, a, a + n))
            a += n
        return layout
    try:
        arr = np.asarray(region, float)
    except ValueError:  # node lists of different lengths
        return None
    if arr.ndim > 2 and arr.shape[-1] == 2:
        shape = arr.shape[:-1]
        return [(None, shape, 0, int(np.prod(shape)))]
    return None


def _region_map(region):
    """AdaptiveMap for an integration region given as limits or node lists,
    a dictionary of them, or an array of limits (pyx:1209-1253)."""
    if isinstance(region, AdaptiveMap):
        return AdaptiveMap(region)
    if isinstance(region, Integrator):
        return AdaptiveMap(region.map)
    if hasattr(region, 'keys'):
        lims = [np.asarray(region[k], float).reshape(-1, 2) for k in region]
        return AdaptiveMap(np.concatenate(lims, axis=0).tolist())
    arr = np.array(region, dtype=object)
    if np.shape(arr.flat[0]) == ():
        grid = np.array(region, float).reshape(-1, 2).tolist()
    else:
        grid = [list(np.asarray(g, float).reshape(-1)) for g in arr.reshape(-1)]
    return AdaptiveMap(grid)


def _as_device(torch, fx):
    if isinstance(fx, torch.Tensor):
        return fx.to(device='cuda', dtype=torch.float64)
    return torch.as_tensor(np.asarray(fx, dtype=float), dtype=torch.float64).cuda()


def _flatten_dict(torch, fx, n, kind, call):
    """A dictionary of batch arrays (the reference's dict-valued batch
    integrands) as one (n, m) array, keys in the integrand's order; the
    layout (key, shape, first, last column) is recorded on the first call."""
    cols, layout, a = [], [], 0
    for k, v in fx.items():
        v = _as_device(torch, v)
        shape = tuple(v.shape[1:]) if kind != 'rbatch' else tuple(v.shape[:-1])
        size = int(np.prod(shape)) if shape else 1
        cols.append(v.reshape(-1, n).T if kind == 'rbatch' else v.reshape(n, -1))
        layout.append((k, shape, a, a + size))
        a += size
    if call.layout is None:
        call.layout = layout
    elif [(k, sh) for k, sh, _, _ in call.layout] != [(k, sh) for k, sh, _, _ in layout]:
        raise ValueError('integrand output changed structure')
    return torch.cat(cols, dim=1)


class Integrator:
    """Adaptive multidimensional Monte Carlo integration (vegas+)."""

    # pyx:1115-1140, plus device options (seed, precision, exact)
    defaults = dict(
        map=None, neval=1000, maxinc_axis=1000, min_neval_batch=100000, max_neval_hcube=50000,
        gpu_pad=False, neval_frac=0.75, max_mem=1e9, nitn=10, alpha=0.5, beta=0.75, adapt=True,
        correlate_integrals=True, minimize_mem=False, adapt_to_errors=False,
        uniform_nstrat=False, rtol=0, atol=0, analyzer=None, ran_array_generator=None,
        sync_ran=True, mpi=False, uses_jac=False, nproc=1,
        seed=None, precision=None, exact=False, deterministic=True,
    )

    def __init__(self, map, **kargs):
        self.neval_hcube_range = None
        self.last_neval = 0
        self._rng_iter = 0
        self._sigf = None
        self.sum_sigf = HUGE
        self._xlayout = _region_layout(map)
        if isinstance(map, Integrator):
            self.map = AdaptiveMap(map.map)
            args = {k: getattr(map, k) for k in Integrator.defaults if k != 'map'}
            self.nstrat = np.array(map.nstrat)
            self._sigf = None if map._sigf is None else map._sigf.clone()
            self.sum_sigf = map.sum_sigf
            # a copy samples independently of its original, as copies that
            # share the reference's global gvar.RNG do: a fresh seed unless
            # seed= is given (pickling keeps the seed and stream position)
            args['seed'] = None
            if 'nstrat' not in kargs:
                kargs['nstrat'] = map.nstrat
                if 'neval' not in kargs:
                    kargs['neval'] = map.neval
        else:
            args = dict(Integrator.defaults)
            del args['map']
            self.map = _region_map(map)
            self.nstrat = np.full(self.map.dim, 0, dtype=np.intp)
        for k, v in args.items():
            setattr(self, k, v)
        args.update(kargs)
        if 'nstrat' in kargs and 'neval' not in kargs and 'neval' in args:
            del args['neval']
        if 'neval' in kargs and 'nstrat' not in kargs and 'nstrat' in args:
            del args['nstrat']
        self._old_neval_given = 'neval' in args
        self.set(args)
        if self.seed is None:
            self.seed = int(np.random.SeedSequence().generate_state(1, np.uint64)[0])

    # ------------------------------------------------------------ settings
    @property
    def sigf(self):
        return np.array([]) if self._sigf is None else self._sigf.cpu().numpy()

    @property
    def dim(self):
        return self.map.dim

    def __reduce__(self):
        odict = {k: getattr(self, k) for k in Integrator.defaults if k not in ('map',)}
        odict['nstrat'] = np.asarray(self.nstrat)
        odict['sigf'] = self.sigf
        # beyond the reference's state: the random stream position and the
        # device's own sum of sigf, so that a restored integrator continues
        # with exactly the points and allocation of the original
        # (the constructor's default settings re-grid the map in passing; the
        # nodes are put back as they were)
        odict['_rng_iter'] = self._rng_iter
        odict['_sum_sigf'] = self.sum_sigf
        odict['_grid'] = self.map.extract_grid()
        odict['_xlayout'] = self._xlayout
        return (Integrator, (self.map,), odict)

    def __setstate__(self, odict):
        rng_iter = odict.pop('_rng_iter', 0)
        sum_sigf = odict.pop('_sum_sigf', None)
        grid = odict.pop('_grid', None)
        xlayout = odict.pop('_xlayout', None)
        self.set(odict)
        self._xlayout = xlayout
        self._rng_iter = rng_iter
        if sum_sigf is not None and self._sigf is not None:
            self.sum_sigf = sum_sigf
        if grid is not None and [len(g) - 1 for g in grid] == [int(n) for n in self.map.ninc]:
            self.map = AdaptiveMap(grid)

    def set(self, ka={}, **kargs):
        """Reset parameters; returns the old values (pyx:1256-1446)."""
        torch = _lib.torch_cuda()
        if kargs:
            kargs.update(ka)
        else:
            kargs = dict(ka)
        old_val = {}
        nstrat = None
        for k in kargs:
            if k == 'map':
                old_val['map'] = self.map
                # the old map carries the old region's structure, so that
                # set(old) restores both
                self.map._xlayout = self._xlayout
                self.map = _region_map(kargs['map'])
                self._xlayout = _region_layout(kargs['map'])
            elif k == 'nstrat':
                if kargs['nstrat'] is None:
                    continue
                old_val['nstrat'] = self.nstrat
                nstrat = np.array(kargs['nstrat'], dtype=np.intp)
            elif k == 'sigf':
                old_val['sigf'] = self.sigf
                s = np.fabs(np.asarray(kargs['sigf'], float))
                self._sigf = torch.as_tensor(s, dtype=torch.float64).cuda()
                self.sum_sigf = float(np.sum(s))
            elif k in Integrator.defaults:
                old_val[k] = getattr(self, k, None)
                v = kargs[k]
                if k in ('neval', 'nitn', 'min_neval_batch', 'max_neval_hcube', 'maxinc_axis', 'nproc'):
                    v = int(v) if v is not None else v
                setattr(self, k, v)
            elif k not in ('nhcube_batch', 'max_nhcube'):
                raise AttributeError('no parameter named "%s"' % str(k))
        if self.nproc not in (1, None) or self.mpi:
            warnings.warn('nproc/mpi are ignored: the iteration runs on one GPU')
        # 2) sanity checks
        if nstrat is not None:
            if len(nstrat) != self.map.dim:
                raise ValueError('nstrat[d] has wrong length: %d not %d' % (len(nstrat), self.map.dim))
            if np.any(nstrat < 1):
                raise ValueError('bad nstrat: ' + str(np.asarray(self.nstrat)))
        if self.neval_frac < 0 or self.neval_frac >= 1:
            raise ValueError('neval_frac = {} but require 0 <= neval_frac < 1'.format(self.neval_frac))
        if 'neval' in old_val and self.neval < 2:
            raise ValueError('neval>2 required, not ' + str(self.neval))
        nf = _geometry.effective_neval_frac(self.neval_frac, self.beta, self.adapt_to_errors)
        # 3) strata per axis
        if nstrat is not None:
            nhcube = int(np.prod(nstrat))
            if 'neval' not in old_val:
                old_val['neval'] = self.neval
                self.neval = int(2. * nhcube / (1. - nf))
            elif self.neval < 2. * nhcube / (1. - nf):
                raise ValueError('neval too small: {} < {}'.format(self.neval, 2. * nhcube / (1. - nf)))
        elif 'neval' in old_val or 'neval_frac' in old_val:
            nstrat = _geometry.auto_nstrat(self.dim, self.neval, nf, self.uniform_nstrat,
                                           self.max_mem, self.minimize_mem)
        else:
            nstrat = np.array(self.nstrat)
        # 4) increments aligned with strata
        nstrat, ninc = _geometry.increments(nstrat, self.neval, self.maxinc_axis, self.adapt_to_errors)
        if not np.all(np.equal(self.map.ninc, ninc)):
            self.map.adapt(ninc=ninc)
        if not np.all(np.equal(self.nstrat, nstrat)):
            if 'sigf' not in old_val:
                old_val['sigf'] = self.sigf
                self._sigf = None
                self.sum_sigf = HUGE
            self.nstrat = nstrat
        # 5) min_neval_hcube
        self.nhcube = int(np.prod(self.nstrat))
        self.min_neval_hcube = _geometry.min_neval_hcube(self.neval, self.nhcube, nf)
        # 6) stratification weights (sigf) live on the device
        if self.beta >= 0 and self.nhcube > 1 and not self.adapt_to_errors and \
                (self._sigf is None or self._sigf.numel() != self.nhcube):
            self._sigf = torch.ones(self.nhcube, dtype=torch.float64, device='cuda')
            self.sum_sigf = float(self.nhcube)
        _geometry.workspace(self.neval, self.dim, self.nhcube, self.min_neval_batch,
                            self.max_neval_hcube, self.max_mem, self.minimize_mem)
        return old_val

    def settings(self, ngrid=0):
        """Summary of the integrator settings (pyx:1448-1590)."""
        nhcube = int(np.prod(self.nstrat))
        neval = nhcube * self.min_neval_hcube if self.beta <= 0 else self.neval
        ans = "Integrator Settings:\n"
        if self.beta > 0 and not self.adapt_to_errors:
            ans += "    %.6g (approx) integrand evaluations in each of %d iterations\n" % (self.neval, self.nitn)
        else:
            ans += "    %.6g integrand evaluations in each of %d iterations\n" % (neval, self.nitn)
        ans += "    number of: strata/axis = %s\n" % np.array2string(
            np.asarray(self.nstrat), max_line_width=80, prefix=29 * ' ')
        ans += "               increments/axis = %s\n" % np.array2string(
            np.asarray(self.map.ninc), max_line_width=80, prefix=33 * ' ')
        ans += "               h-cubes = %.6g  processors = %d\n" % (nhcube, 1)
        max_neval_hcube = max(self.max_neval_hcube, self.min_neval_hcube)
        ans += "               evaluations/batch >= %.2g\n" % (float(self.min_neval_batch),)
        ans += "               %d <= evaluations/h-cube <= %.2g\n" % (int(self.min_neval_hcube),
                                                                     float(max_neval_hcube))
        ans += "    minimize_mem = %s  adapt_to_errors = %s  adapt = %s\n" % (
            str(self.minimize_mem), str(self.adapt_to_errors), str(self.adapt))
        ans += "    accuracy: relative = %g  absolute = %g\n" % (self.rtol, self.atol)
        if not self.adapt:
            ans += "    damping: alpha = %g  beta= %g\n\n" % (0., 0.)
        elif self.adapt_to_errors:
            ans += "    damping: alpha = %g  beta= %g\n\n" % (self.alpha, 0.)
        else:
            ans += "    damping: alpha = %g  beta= %g\n\n" % (self.alpha, self.beta)
        ans += self._limits_table()
        if ngrid > 0:
            ans += '\n' + self.map.settings(ngrid=ngrid)
        return ans

    def _limits_table(self):
        """The integration-limits table of settings() (pyx:1508-1586): a
        key/index column for dictionary or multi-dimensional regions (the
        key printed on a variable's first index only), two columns side by
        side beyond 20 axes."""
        limits = ['({:.5}, {:.5})'.format(*lim) for lim in self.map.region()]

        def idx_str(idx):
            return ','.join(str(i) for i in idx)

        if self._xlayout is None:
            keys = None
        else:
            keys = []
            for k, shape, a, b in self._xlayout:
                if k is None:
                    keys += [idx_str(idx) for idx in np.ndindex(shape)]
                elif shape == ():
                    keys.append(str(k))
                else:
                    for j, idx in enumerate(np.ndindex(shape)):
                        keys.append((str(k) + ' ' if j == 0 else '') + idx_str(idx))
        entries = [(None if keys is None else keys[a], str(a), lim) for a, lim in enumerate(limits)]
        ncol = 1 if self.map.dim <= 20 else 2
        nl = -(-len(entries) // ncol)
        cols = [entries[c * nl:(c + 1) * nl] for c in range(ncol)]
        w0 = None if keys is None else max(max(len(e[0]) for e in entries), len('key/index'))
        w1 = max(max(len(e[1]) for e in entries), len('axis'))
        w2 = max(max(len(e[2]) for e in entries), len('integration limits'))

        def fmt(e0, e1, e2):
            s = '' if w0 is None else '{0:>{1}}    '.format(e0, w0)
            return s + '{0:>{1}}    {2:>{3}}'.format(e1, w1, e2, w2)

        tables = []
        for col in cols:
            head = fmt('key/index', 'axis', 'integration limits')
            tables.append([head, len(head) * '-'] + [fmt(*e) for e in col])
        lines = []
        for i in range(len(tables[0])):
            lines.append('  '.join(t[i] for t in tables if i < len(t)))
        return 4 * ' ' + ('\n' + 4 * ' ').join(lines) + '\n'

    # --------------------------------------------------------- device plumbing
    def _precision(self, device_fcn):
        p = self.precision
        if p is None:
            p = 'f32' if device_fcn else 'f64'
        if p not in ('f32', 'f64'):
            raise ValueError("precision must be 'f32' or 'f64'")
        if p == 'f32' and self.ran_array_generator is not None:
            raise ValueError("ran_array_generator gives float64 uniforms: use precision='f64'")
        return _lib.PREC_F32 if p == 'f32' else _lib.PREC_F64

    def _uniforms(self, torch, cfg, total):
        """Device buffer [total, dim] for the caller's uniforms of one
        iteration (ran_array_generator, pyx:1676-1680), attached to cfg; None
        when the samples come from the device Philox stream."""
        if self.ran_array_generator is None:
            return None
        u = torch.empty((max(int(total), 1), self.dim), dtype=torch.float64, device='cuda')
        cfg.uniforms = u.data_ptr()
        return u

    def _draw(self, torch, u, p0, p1):
        """ran_array_generator((n, dim)) for the batch of points [p0, p1)
        (pyx:1732), into rows p0..p1 of the iteration's uniform buffer."""
        if u is None:
            return
        n = p1 - p0
        r = self.ran_array_generator((n, self.dim))
        r = torch.as_tensor(np.asarray(r, dtype=np.float64) if not torch.is_tensor(r) else r,
                            dtype=torch.float64)
        if tuple(r.shape) != (n, self.dim):
            raise ValueError('ran_array_generator((%d, %d)) returned shape %s' % (n, self.dim, tuple(r.shape)))
        u[p0:p1].copy_(r)

    def _config(self, nitn, iteration, prec):
        c = _lib.VgxConfig()
        c.dim = self.dim
        c.ld = self.map.ld
        for d in range(self.dim):
            c.ninc[d] = int(self.map.ninc[d])
            c.nstrat[d] = int(self.nstrat[d])
        c.nhcube = self.nhcube
        c.neval = int(self.neval)
        c.neval_frac = float(self.neval_frac)
        c.alpha = float(self.alpha)
        c.beta = float(self.beta)
        c.adapt = int(bool(self.adapt))
        c.adapt_to_errors = int(bool(self.adapt_to_errors))
        c.min_neval_hcube = int(self.min_neval_hcube)
        c.max_neval_hcube = int(self.max_neval_hcube)
        c.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        c.iteration = int(iteration) & 0xFFFFFFFF
        c.nitn = int(nitn)
        c.precision = prec
        c.flags = (_lib.FLAG_EXACT_ADAPT | _lib.FLAG_EXACT_DIV) if self.exact else 0
        if not self.deterministic:
            c.flags |= _lib.FLAG_UNORDERED_HIST
        c.weighted = int(bool(self.adapt))
        return c

    def _state(self, torch, nitn):
        """Device state of one call: the training histogram (initialised by
        the library at the start of a run) and one small float64 block,
        filled from the host in one copy and read back in one copy, that
        holds sum_sigf, the RAvg summary, the status word and the
        per-iteration results (int64 counts as views of the same block)."""
        dev = 'cuda'
        f64 = torch.float64
        if self._sigf is None:
            self._sigf = torch.ones(max(self.nhcube, 1), dtype=f64, device=dev)
        nb = self.dim * (self.map.ld - 1)
        hist = torch.empty(2 * nb, dtype=f64, device=dev)
        h = np.zeros(6 + 5 * nitn)
        h[0] = self.sum_sigf
        head = torch.from_numpy(h).to(dev)
        o = 6 + 2 * nitn
        bufs = dict(
            head=head,
            hist=hist,
            sum_sigf=head[0:1],
            ravg=head[1:5],
            status=head[5:6].view(torch.int32),
            itn_mean=head[6:6 + nitn],
            itn_var=head[6 + nitn:o],
            itn_neval=head[o:o + nitn].view(torch.int64),
            itn_nhrange=head[o + nitn:o + 3 * nitn].view(torch.int64),
            sum_f=hist[:nb],
            n_f=hist[nb:],
        )
        self._train_hist = hist
        st = _lib.VgxState()
        st.grid = self.map._grid.data_ptr()
        st.sigf = self._sigf.data_ptr()
        for k in ('sum_sigf', 'ravg', 'status', 'itn_mean', 'itn_var', 'itn_neval', 'itn_nhrange',
                  'sum_f', 'n_f'):
            setattr(st, k, bufs[k].data_ptr())
        return st, bufs

    def _read_back(self, bufs):
        # one synchronising copy of everything the host needs
        h = bufs['head'].cpu().numpy()
        nitn = (len(h) - 6) // 5
        o = 6 + 2 * nitn
        status = int(h[5:6].view(np.int32)[0])
        res = dict(sum_sigf=h[0:1], ravg=h[1:5], itn_mean=h[6:6 + nitn], itn_var=h[6 + nitn:o],
                   itn_neval=h[o:o + nitn].view(np.int64), itn_nhrange=h[o + nitn:o + 3 * nitn].view(np.int64))
        return status, res

    # ---------------------------------------------------------- integration
    def __call__(self, fcn, save=None, saveall=None, **kargs):
        """Integrate ``fcn`` (pyx:1912-2230); returns an :class:`RAvg`."""
        if kargs:
            self.set(kargs)
        torch = _lib.torch_cuda()
        device_fcn = isinstance(fcn, DeviceIntegrand)
        if device_fcn and fcn.dim != self.dim:
            raise ValueError('integrand dimension %d != integrator dimension %d' % (fcn.dim, self.dim))
        if self.ran_array_generator is not None:
            # the caller's uniforms feed the staged iteration; a device
            # integrand is then called through its batch form
            device_fcn = False
        prec = self._precision(device_fcn)
        std = None if device_fcn else self._std_integrand(fcn)
        layout = None if device_fcn else self._probe_layout(torch, std)
        # scalar integrands give an RAvg, arrays an RAvgArray of their shape,
        # dictionaries an RAvgDict (pyx:2058-2064)
        if layout is None or (len(layout) == 1 and layout[0][0] is None and layout[0][1] == ()):
            result = RAvg(weighted=self.adapt)
            if std is not None:
                std.multi = False
        elif len(layout) == 1 and layout[0][0] is None:
            result = RAvgArray(layout[0][1], weighted=self.adapt)
            std.multi = True
        else:
            result = RAvgDict(layout, weighted=self.adapt)
            std.multi = True
        per_itn = bool(self.rtol or self.atol or self.analyzer is not None or save or saveall)
        if device_fcn and not per_itn:
            self._run_device(torch, fcn, self.nitn, prec, result)
        else:
            for itn in range(self.nitn):
                if self.analyzer is not None:
                    self.analyzer.begin(itn, self)
                if device_fcn:
                    self._run_device(torch, fcn, 1, prec, result)
                else:
                    self._run_staged(torch, std, prec, result)
                if self.analyzer is not None:
                    self.analyzer.end(result.itn_results[-1], result)
                if save is not None:
                    _dump(save, result)
                if saveall is not None:
                    _dump(saveall, (result, self))
                if result.converged(self.rtol, self.atol):
                    break
        self.map.clear()
        return result

    def _finish(self, status, res, n, result):
        if status == _lib.VGX_ENAN:
            raise ValueError('integrand evaluates to nan')
        if status:
            _lib.check(status, 'integrate')
        result.add_all([GVar(float(res['itn_mean'][k]), math.sqrt(max(float(res['itn_var'][k]), 0.0))
                             if res['itn_var'][k] >= 0 else float('nan')) for k in range(n)])
        result.sum_neval += int(np.sum(res['itn_neval'][:n]))
        self.last_neval = int(res['itn_neval'][n - 1])
        self.neval_hcube_range = np.array(res['itn_nhrange'][2 * (n - 1):2 * n])
        if self.beta > 0 and not self.adapt_to_errors and self.adapt:
            self.sum_sigf = float(res['sum_sigf'][0])
        self._last_device_ravg = res['ravg']

    def _last_training(self):
        """The training histogram (sum_f, n_f), each [dim, max(ninc)], of the
        last iteration run on the device, as the adapt after it read it
        (before AdaptiveMap.clear in the reference, pyx:2216-2217, 596-599)."""
        hist = self._train_hist.cpu().numpy()
        nb = len(hist) // 2
        return hist[:nb].reshape(self.dim, -1), hist[nb:].reshape(self.dim, -1)

    def _run_device(self, torch, fcn, nitn, prec, result):
        st, bufs = self._state(torch, nitn)
        f, keep = fcn.vgx_integrand(torch, 'cuda')
        cfg = self._config(nitn, self._rng_iter, prec)
        _lib.check(_lib.lib().vgx_integrate(_lib.ctx(), C.byref(cfg), C.byref(f), C.byref(st),
                                            _lib.stream_ptr()), 'vgx_integrate')
        status, res = self._read_back(bufs)
        self._rng_iter += nitn
        self._finish(status, res, nitn, result)

    def _std_integrand(self, fcn):
        """Normalise fcn to f(x[n, dim] device float64) -> f[n] (pyx:2959-3383
        for scalar integrands)."""
        torch = _lib.torch_cuda()
        kind = getattr(fcn, 'fcntype', 'scalar')
        on_gpu = bool(getattr(fcn, 'vgx_torch', False))
        uses_jac = self.uses_jac

        xl = self._xlayout

        def shaped(v, mode):
            # flat axes -> the region's structure: a dict per key or one array
            # (batch index first, last, or none for a single point)
            if xl is None or v is None:
                return v
            if mode == 'point':
                out = {k: (v[a] if shape == () else v[a:b].reshape(shape)) for k, shape, a, b in xl}
            elif mode == 'lbatch':
                out = {k: (v[:, a] if shape == () else v[:, a:b].reshape((v.shape[0],) + shape))
                       for k, shape, a, b in xl}
            else:
                out = {k: (v[a] if shape == () else v[a:b].reshape(shape + (v.shape[-1],))) for k, shape, a, b in xl}
            return out[None] if xl[0][0] is None else out

        def call(x, jac):
            if kind == 'rbatch':
                xa, ja = x.T, (None if jac is None else jac.T)
            elif kind == 'lbatch':
                xa, ja = x, jac
            else:
                xa, ja = x, jac
            bmode = 'rbatch' if kind == 'rbatch' else 'lbatch'
            if on_gpu:
                xs, js = shaped(xa, bmode), shaped(ja, bmode)
                fx = fcn(xs) if js is None else fcn(xs, jac=js)
            else:
                xn = xa.cpu().numpy()
                jn = None if ja is None else ja.cpu().numpy()
                if kind in ('lbatch', 'rbatch'):
                    xs, js = shaped(xn, bmode), shaped(jn, bmode)
                    fx = fcn(xs) if js is None else fcn(xs, jac=js)
                else:
                    vals = [fcn(shaped(xi, 'point')) if jn is None else fcn(shaped(xi, 'point'), jac=shaped(jn[i], 'point'))
                            for i, xi in enumerate(xn)]
                    if vals and isinstance(vals[0], dict):
                        fx = {k: np.array([v[k] for v in vals], dtype=float) for k in vals[0]}
                    else:
                        fx = np.array(vals, dtype=float)
            n = x.shape[0]
            if isinstance(fx, dict):
                fx = _flatten_dict(torch, fx, n, kind, call)
            else:
                fx = _as_device(torch, fx)
                if call.layout is None:
                    shape = tuple(fx.shape[1:]) if kind != 'rbatch' else tuple(fx.shape[:-1])
                    call.layout = [(None, shape, 0, int(np.prod(shape)) if shape else 1)]
                fx = fx.reshape(-1, n).T if kind == 'rbatch' else fx.reshape(n, -1)
            return fx.contiguous()

        call.uses_jac = uses_jac
        call.layout = None
        return call

    def _probe_layout(self, torch, std):
        """Output structure of a Python integrand: one call on a batch of one
        point (the centre of the y box) with jac = 1, as the reference's
        VegasIntegrand evaluates fcn at xsample to learn its shape and size
        (pyx:3010-3087)."""
        y = torch.full((1, self.dim), 0.5, dtype=torch.float64, device='cuda')
        x, _ = self.map.map(y)
        std(x, torch.ones_like(y) if std.uses_jac else None)
        return std.layout

    def _run_staged(self, torch, std, prec, result):
        L = _lib.lib()
        st, bufs = self._state(torch, 1)
        cfg = self._config(1, self._rng_iter, prec)
        nh = torch.empty(self.nhcube, dtype=torch.int32, device='cuda')
        total = C.c_int64()
        cx, sp = _lib.ctx(), _lib.stream_ptr()
        m = std.layout[-1][3] if std.layout else 1
        multi = bool(getattr(std, 'multi', m > 1))
        corr = 1 if self.correlate_integrals else 0
        if multi:
            # per iteration: mean[m] and the covariance lower triangle (or
            # the variances), double-double pairs
            npairs = m * (m + 1) // 2 if corr else m
            acc = torch.zeros(2 * (m + npairs), dtype=torch.float64, device='cuda')
            cstat = torch.zeros(1, dtype=torch.int32, device='cuda')
        _lib.check(L.vgx_iter_begin(cx, C.byref(cfg), C.byref(st), 0, _lib.ptr(nh), C.byref(total),
                                    sp), 'iter_begin')
        ubuf = self._uniforms(torch, cfg, total.value)
        # NaN in any component: flagged on the device batch by batch and read
        # once, before the iteration closes (no adapt), instead of a host
        # synchronisation per batch (pyx:2133-2134 raises at the batch)
        nan_seen = torch.zeros((), dtype=torch.bool, device='cuda')
        for h0, h1, p0, p1 in _geometry.batches(nh.cpu().numpy(), self.min_neval_batch):
            n = p1 - p0
            self._draw(torch, ubuf, p0, p1)
            x = torch.empty((n, self.dim), dtype=torch.float64, device='cuda')
            y = torch.empty((n, self.dim), dtype=torch.float64, device='cuda') if std.uses_jac else None
            w = torch.empty(n, dtype=torch.float64, device='cuda') if multi else None
            _lib.check(L.vgx_iter_sample(cx, C.byref(cfg), C.byref(st), 0, h0, h1, p0, p1, _lib.ptr(x),
                                         _lib.ptr(y), _lib.ptr(w), None, sp), 'iter_sample')
            jac = self.map.jac1d(y) if y is not None else None
            f = std(x, jac)
            if f.shape[1] != m:
                raise ValueError('integrand output changed shape (%d values, not %d)' % (f.shape[1], m))
            nan_seen |= torch.isnan(f).any()
            # component 0 steers sigf and the training data (pyx:2173, 2185)
            f0 = f[:, 0].contiguous()
            _lib.check(L.vgx_iter_accumulate(cx, C.byref(cfg), C.byref(st), 0, h0, h1, p0, p1,
                                             _lib.ptr(f0), sp), 'iter_accumulate')
            if multi:
                _lib.check(L.vgx_iter_accumulate_cov(cx, C.byref(cfg), h0, h1, p0, p1, _lib.ptr(f), _lib.ptr(w),
                                                     m, corr, _lib.ptr(acc), _lib.ptr(cstat), sp),
                           'iter_accumulate_cov')
        if bool(nan_seen):
            raise ValueError('integrand evaluates to nan')
        _lib.check(L.vgx_iter_end(cx, C.byref(cfg), C.byref(st), 0, sp), 'iter_end')
        status, res = self._read_back(bufs)
        self._rng_iter += 1
        if not multi:
            self._finish(status, res, 1, result)
            return
        if int(cstat.item()):
            _lib.check(int(cstat.item()), 'iter_accumulate_cov')
        scalar = RAvg(weighted=self.adapt)
        self._finish(status, res, 1, scalar)
        a = acc.cpu().numpy()
        mean = a[0:2 * m:2] + a[1:2 * m:2]
        cv = a[2 * m::2] + a[2 * m + 1::2]
        cov = np.zeros((m, m))
        if corr:
            il = np.tril_indices(m)
            cov[il] = cv
            cov = cov + np.tril(cov, -1).T
        else:
            cov[np.diag_indices(m)] = cv
        result.add(GVarArray(mean, cov, (m,)))
        result.sum_neval += int(res['itn_neval'][0])
        self._last_scalar = scalar

    # -------------------------------------------------------- sampler APIs
    def _device_batches(self, yield_hcube=False, yield_y=False):
        """Device tensors (x, [y,] wgt[, hcube]) of one iteration's points,
        in the batches _random_batch cuts (pyx:1636-1765)."""
        torch = _lib.torch_cuda()
        L = _lib.lib()
        st, bufs = self._state(torch, 1)
        cfg = self._config(1, self._rng_iter, self._precision(False))
        nh = torch.empty(self.nhcube, dtype=torch.int32, device='cuda')
        total = C.c_int64()
        cx, sp = _lib.ctx(), _lib.stream_ptr()
        _lib.check(L.vgx_iter_begin(cx, C.byref(cfg), C.byref(st), 0, _lib.ptr(nh), C.byref(total), sp),
                   'iter_begin')
        self._rng_iter += 1
        nhh = nh.cpu().numpy()
        self.last_neval = int(total.value)
        ubuf = self._uniforms(torch, cfg, total.value)
        for h0, h1, p0, p1 in _geometry.batches(nhh, self.min_neval_batch):
            n = p1 - p0
            self._draw(torch, ubuf, p0, p1)
            x = torch.empty((n, self.dim), dtype=torch.float64, device='cuda')
            y = torch.empty_like(x) if yield_y else None
            w = torch.empty(n, dtype=torch.float64, device='cuda')
            hc = torch.empty(n, dtype=torch.int64, device='cuda') if yield_hcube else None
            _lib.check(L.vgx_iter_sample(cx, C.byref(cfg), C.byref(st), 0, h0, h1, p0, p1, _lib.ptr(x),
                                         _lib.ptr(y), _lib.ptr(w), _lib.ptr(hc), sp), 'iter_sample')
            ans = (x,)
            if yield_y:
                ans += (y,)
            ans += (w,)
            if yield_hcube:
                ans += (hc,)
            yield ans

    def random_batch(self, yield_hcube=False, yield_y=False):
        """Batches of integration points and weights for one iteration
        (pyx:1601-1634, 1636-1765): yields (x, [y,] wgt[, hcube])."""
        for t in self._device_batches(yield_hcube, yield_y):
            yield tuple(a.cpu().numpy() for a in t)

    random_vec = random_batch

    def random(self, yield_hcube=False, yield_y=False):
        """Single points and weights (pyx:1770-1814)."""
        for t in self.random_batch(yield_hcube=yield_hcube, yield_y=yield_y):
            for i in range(t[0].shape[0]):
                yield tuple(ti[i] for ti in t)

    def sample(self, nbatch=None, mode='rbatch'):
        """Sample of weights and points covering whole iterations (pyx:1816-1863)."""
        neval = self.last_neval if self.last_neval > 0 else self.neval
        nbatch = neval if nbatch is None else int(nbatch)
        nit = nbatch // neval
        if nit * neval < nbatch:
            nit += 1
        samples, wgts = [], []
        for _ in range(nit):
            for x, w in self.random_batch():
                samples.append(x)
                wgts.append(w)
        samples = np.concatenate(samples, axis=0)
        wgts = np.concatenate(wgts) / nit
        xl = self._xlayout
        if xl is not None and xl[0][0] is not None:
            # dictionary region: one array per key, batch index last (rbatch)
            # or first (lbatch), pyx:1851-1855
            n = samples.shape[0]
            out = {}
            for k, shape, a, b in xl:
                v = samples[:, a:b].reshape((n,) + shape)
                out[k] = np.moveaxis(v, 0, -1) if mode == 'rbatch' else v
            return wgts, out
        if xl is not None:
            shape = xl[0][1]
            samples = samples.reshape((-1,) + shape)
            return wgts, (np.moveaxis(samples, 0, -1) if mode == 'rbatch' else samples)
        if mode == 'rbatch':
            samples = samples.T
        return wgts, samples


def _dump(target, obj):
    if isinstance(target, str):
        with open(target, 'wb') as f:
            pickle.dump(obj, f)
    else:
        pickle.dump(obj, target)


class reporter:
    """Analyzer printing a per-iteration report (pyx:2232-2269)."""

    def __init__(self, ngrid=0):
        import time
        self.ngrid = ngrid
        self.clock = time.perf_counter

    def begin(self, itn, integrator):
        self.integrator = integrator
        self.itn = itn
        self.t0 = self.clock()
        if itn == 0:
            print(integrator.settings())

    def end(self, itn_ans, ans):
        print("    itn %2d: %s\n all itn's: %s" % (self.itn + 1, itn_ans, ans))
        print('    neval = %s  neval/h-cube = %s\n    chi2/dof = %.2f  Q = %.2f  time = %.2f' % (
            format(self.integrator.last_neval, '.6g'), tuple(self.integrator.neval_hcube_range),
            ans.chi2 / ans.dof if ans.dof > 0 else 0, ans.Q if ans.dof > 0 else 1.,
            self.clock() - self.t0))
        print(self.integrator.map.settings(ngrid=self.ngrid))
        print('')
