"""Device integrands: evaluated inside the fused sampler kernel.

A Python callable cannot run inside a GPU kernel, so integrand families
used at scale are compiled into libvgx as functors selected by id plus a
parameter block (include/vgx.h, vgx_integrand).  Each class here describes
one family; passing an instance to :class:`vegas_amd.Integrator` runs the
whole iteration on the device.  ``__call__`` evaluates the same function on
host numpy arrays for convenience (it is not used by the integrator).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib


class DeviceIntegrand:
    kind = None
    fcntype = 'lbatch'

    def _device_arrays(self, torch, device):
        key = (str(device),)
        cache = self.__dict__.setdefault('_dev', {})
        if key not in cache:
            cache[key] = tuple(torch.as_tensor(a, dtype=torch.float64, device=device)
                               for a in self._param_arrays())
        return cache[key]

    def vgx_integrand(self, torch, device):
        A, B, c = self._device_arrays(torch, device)
        f = _lib.VgxIntegrand()
        f.kind = self.kind
        f.ncomp = int(A.numel())
        f.A, f.B, f.c = A.data_ptr(), B.data_ptr(), c.data_ptr()
        return f, (A, B, c)


class GaussianMixture(DeviceIntegrand):
    """f(x) = sum_k A[k] exp(-B[k] |x - centers[k]|^2)."""
    kind = _lib.FCN_GAUSSMIX

    def __init__(self, A, B, centers):
        self.A = np.atleast_1d(np.asarray(A, np.float64))
        self.B = np.atleast_1d(np.asarray(B, np.float64))
        self.centers = np.atleast_2d(np.asarray(centers, np.float64))
        self.ncomp, self.dim = self.centers.shape
        if self.A.shape != (self.ncomp,) or self.B.shape != (self.ncomp,):
            raise ValueError('A, B need one entry per centre')

    def _param_arrays(self):
        return self.A, self.B, np.ascontiguousarray(self.centers).reshape(-1)

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        r2 = ((x[:, None, :] - self.centers[None, :, :]) ** 2).sum(-1)
        return (self.A[None, :] * np.exp(-self.B[None, :] * r2)).sum(-1)

    @classmethod
    def isotropic(cls, centers, sigma, weights=None):
        """Sum of normalised isotropic Gaussians (each integrates to its
        weight over R^d); weights default to 1/ncomp."""
        centers = np.atleast_2d(np.asarray(centers, np.float64))
        nc, dim = centers.shape
        sigma = np.broadcast_to(np.asarray(sigma, np.float64), (nc,))
        w = np.full(nc, 1.0 / nc) if weights is None else np.asarray(weights, np.float64)
        A = w * (2 * math.pi * sigma ** 2) ** (-dim / 2.)
        B = 1.0 / (2 * sigma ** 2)
        return cls(A, B, centers)


class DiagonalRidge(DeviceIntegrand):
    """Mixture with centres t[k]*(1,...,1) on the main diagonal:
    f(x) = sum_k A[k] exp(-B[k] |x - t[k] 1|^2), evaluated through
    |x - t 1|^2 = sum_d (x_d - xbar)^2 + dim (xbar - t)^2."""
    kind = _lib.FCN_DIAGRIDGE

    def __init__(self, A, B, t, dim):
        self.A = np.atleast_1d(np.asarray(A, np.float64))
        self.B = np.atleast_1d(np.asarray(B, np.float64))
        self.t = np.atleast_1d(np.asarray(t, np.float64))
        self.dim = int(dim)
        self.ncomp = len(self.t)

    def _param_arrays(self):
        return self.A, self.B, self.t

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        r2 = ((x[:, None, :] - self.t[None, :, None]) ** 2).sum(-1)
        return (self.A[None, :] * np.exp(-self.B[None, :] * r2)).sum(-1)

    def as_mixture(self):
        return GaussianMixture(self.A, self.B, self.t[:, None] * np.ones(self.dim)[None, :])
