"""The reference's batch-integrand convention (src/vegas/_vegas.pyx:3385-3473).

``lbatchintegrand``/``rbatchintegrand`` tag a callable with ``fcntype``
'lbatch' (batch index first) or 'rbatch' (batch index last), exactly as the
reference does; ``LBatchIntegrand``/``RBatchIntegrand`` are the equivalent
base/wrapper classes.  Integrands so tagged receive numpy float64 arrays, as
in the reference.

``gpu_lbatchintegrand``/``gpu_rbatchintegrand`` (no reference counterpart)
additionally mark the integrand as taking torch tensors on the GPU: ``x`` is a
float64 device tensor, valid only during the call, and the integrand returns a
device tensor -- no host round trip.
"""
from __future__ import annotations


class LBatchIntegrand(object):
    """Wrapper for lbatch integrands (pyx:3386-3406)."""

    def __init__(self, fcn=None):
        self.fcn = self if fcn is None else fcn

    @property
    def fcntype(self):
        return 'lbatch'

    def __call__(self, *args, **kargs):
        return self.fcn(*args, **kargs)

    def __getattr__(self, attr):
        if attr == 'fcn':
            raise AttributeError(attr)
        return getattr(self.fcn, attr)


class RBatchIntegrand(object):
    """Same as :class:`LBatchIntegrand` with the batch index on the right."""

    def __init__(self, fcn=None):
        self.fcn = self if fcn is None else fcn

    @property
    def fcntype(self):
        return 'rbatch'

    def __call__(self, *args, **kargs):
        return self.fcn(*args, **kargs)

    def __getattr__(self, attr):
        if attr == 'fcn':
            raise AttributeError(attr)
        return getattr(self.fcn, attr)


def lbatchintegrand(f):
    """Decorator for batch integrands ``f(x[i, d]) -> f[i]`` (pyx:3408-3438)."""
    try:
        f.fcntype = 'lbatch'
        return f
    except Exception:
        return LBatchIntegrand(f)


def rbatchintegrand(f):
    """Decorator for batch integrands ``f(x[d, i]) -> f[i]`` (pyx:3457-3463)."""
    try:
        f.fcntype = 'rbatch'
        return f
    except Exception:
        return RBatchIntegrand(f)


def _gpu(f, kind):
    try:
        f.fcntype = kind
        f.vgx_torch = True
        return f
    except Exception:
        w = LBatchIntegrand(f) if kind == 'lbatch' else RBatchIntegrand(f)
        w.__dict__['vgx_torch'] = True
        return w


def gpu_lbatchintegrand(f):
    """lbatch integrand evaluated on torch device tensors."""
    return _gpu(f, 'lbatch')


def gpu_rbatchintegrand(f):
    """rbatch integrand evaluated on torch device tensors."""
    return _gpu(f, 'rbatch')


# legacy names (pyx:3466-3473)
batchintegrand = lbatchintegrand
BatchIntegrand = LBatchIntegrand
vecintegrand = batchintegrand
MPIintegrand = batchintegrand


class VecIntegrand(LBatchIntegrand):
    pass
