"""Integrator geometry: strata, increments and per-hcube minimum.

Host logic restating ``Integrator.set`` steps 2-6 of gplepage/vegas 6.4.1
(src/vegas/_vegas.pyx:1321-1446).  Pure integer/float arithmetic with the
reference's ``int()`` truncations, ``**`` and ``numpy.log``; it runs on the
CPU and touches no device state.
"""
from __future__ import annotations

import numpy as np


def effective_neval_frac(neval_frac, beta, adapt_to_errors):
    """pyx:1331: stratification share used for geometry."""
    return 0 if (beta == 0 or adapt_to_errors) else neval_frac


def auto_nstrat(dim, neval, nf, uniform_nstrat, max_mem, minimize_mem):
    """pyx:1346-1366: strata per axis from neval."""
    ns = int(abs((1 - nf) * neval / 2.) ** (1. / dim))
    if ns < 1:
        ns = 1
    d = int((np.log((1 - nf) * neval / 2.) - dim * np.log(ns)) / np.log(1 + 1. / ns))
    if ((ns + 1) ** d * ns ** (dim - d)) > max_mem and not minimize_mem:
        raise MemoryError("work arrays larger than max_mem; set minimize_mem=True "
                          "(and install h5py module) or increase max_mem")
    if uniform_nstrat:
        d = 0
    nstrat = np.empty(dim, np.intp)
    nstrat[:d] = ns + 1
    nstrat[d:] = ns
    return nstrat


def increments(nstrat, neval, maxinc_axis, adapt_to_errors):
    """pyx:1371-1386: increments per axis aligned with the strata.  May
    coarsen nstrat (returned) when it exceeds maxinc_axis."""
    nstrat = np.array(nstrat, np.intp)
    dim = len(nstrat)
    if adapt_to_errors:
        return nstrat, nstrat.copy()
    ni = min(int(neval / 10.), maxinc_axis)
    ninc = np.empty(dim, np.intp)
    for d in range(dim):
        if ni >= nstrat[d]:
            ninc[d] = int(ni / nstrat[d]) * nstrat[d]
        elif nstrat[d] <= maxinc_axis:
            ninc[d] = nstrat[d]
        else:
            nstrat[d] = int(nstrat[d] / ni) * ni
            ninc[d] = ni
    return nstrat, ninc


def min_neval_hcube(neval, nhcube, nf):
    """pyx:1400-1407."""
    if nhcube == 1:
        m = int(neval)
    else:
        m = int((1 - nf) * neval / nhcube)
    return max(m, 2)


def workspace(neval, dim, nhcube, min_neval_batch, max_neval_hcube, max_mem, minimize_mem):
    """pyx:1437-1441: the reference's batch buffer size and memory check."""
    ws = min_neval_batch + max_neval_hcube
    if ws > neval:
        ws = neval + 1
    if (3 * dim + 3) * ws + (0 if minimize_mem else nhcube) > max_mem:
        raise MemoryError("work arrays larger than max_mem; reduce min_neval_batch or "
                          "max_neval_hcube (or increase max_mem)")
    return ws


def compute(dim, neval=None, nstrat=None, neval_frac=0.75, beta=0.75, adapt_to_errors=False,
            uniform_nstrat=False, maxinc_axis=1000, max_mem=1e9, minimize_mem=False,
            min_neval_batch=100000, max_neval_hcube=50000):
    """Geometry of a fresh Integrator (neval and/or nstrat given)."""
    if neval_frac < 0 or neval_frac >= 1:
        raise ValueError('neval_frac = {} but require 0 <= neval_frac < 1'.format(neval_frac))
    nf = effective_neval_frac(neval_frac, beta, adapt_to_errors)
    if nstrat is not None:
        nstrat = np.array(nstrat, np.intp)
        if len(nstrat) != dim or np.any(nstrat < 1):
            raise ValueError('bad nstrat = %s' % str(nstrat))
        nhcube = int(np.prod(nstrat))
        if neval is None:
            neval = int(2. * nhcube / (1. - nf))
        elif neval < 2. * nhcube / (1. - nf):
            raise ValueError('neval too small: {} < {}'.format(neval, 2. * nhcube / (1. - nf)))
    else:
        if neval is None:
            raise ValueError('need neval or nstrat')
        if neval < 2:
            raise ValueError('neval>2 required, not ' + str(neval))
        nstrat = auto_nstrat(dim, neval, nf, uniform_nstrat, max_mem, minimize_mem)
    neval = int(neval)
    nstrat, ninc = increments(nstrat, neval, maxinc_axis, adapt_to_errors)
    nhcube = int(np.prod(nstrat))
    mnh = min_neval_hcube(neval, nhcube, nf)
    ws = workspace(neval, dim, nhcube, min_neval_batch, max_neval_hcube, max_mem, minimize_mem)
    return dict(nstrat=nstrat, ninc=ninc, nhcube=nhcube, neval=neval, min_neval_hcube=mnh,
                workspace=ws)


def batches(n_h, min_neval_batch):
    """Batch boundaries of ``_random_batch`` (pyx:1692-1710): a batch closes
    at the first hcube where the running count reaches min_neval_batch, or at
    the last hcube.  Returns a list of (h0, h1, p0, p1)."""
    n_h = np.asarray(n_h, np.int64)
    out = []
    h0 = 0
    p0 = 0
    nh = len(n_h)
    csum = np.cumsum(n_h)
    while h0 < nh:
        base = csum[h0 - 1] if h0 > 0 else 0
        # first h >= h0 with csum[h] - base >= min_neval_batch
        h = int(np.searchsorted(csum, base + min_neval_batch, side='left'))
        h1 = min(max(h, h0) + 1, nh)
        p1 = int(csum[h1 - 1])
        out.append((h0, h1, p0, p1))
        h0, p0 = h1, p1
    return out
