"""vegas_amd -- the vegas+ adaptive Monte Carlo integrator on AMD MI355X.

The call surface is gplepage/vegas' (``Integrator``, ``AdaptiveMap``,
``RAvg``, ``lbatchintegrand``/``rbatchintegrand``); the hot path runs as
hand-written HIP kernels for gfx950 in ``libvgx.so`` (C-ABI in include/vgx.h).
"""
from ._ravg import GVar, GVarArray, RAvg, RAvgArray, RAvgDict, gammaQ, ravg
from .batchintegrand import (BatchIntegrand, LBatchIntegrand, MPIintegrand, RBatchIntegrand,
                             VecIntegrand, batchintegrand, gpu_lbatchintegrand,
                             gpu_rbatchintegrand, lbatchintegrand, rbatchintegrand, vecintegrand)
from .integrands import DeviceIntegrand, DiagonalRidge, GaussianMixture

__version__ = "0.1.0"


def __getattr__(name):
    # device-backed classes import torch lazily
    if name == 'AdaptiveMap':
        from .adaptive_map import AdaptiveMap
        return AdaptiveMap
    if name in ('Integrator', 'reporter'):
        from . import integrator
        return getattr(integrator, name)
    if name in ('PDFIntegrator', 'PDFEV', 'PDFEVArray', 'PDFEVDict', 'PDFStatistics', 'restratify'):
        from . import pdf
        return getattr(pdf, name)
    raise AttributeError(name)
