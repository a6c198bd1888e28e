"""Benchmark workloads: the integrands and run parameters of the measured
workloads (exact integral 1 unless stated)."""
from __future__ import annotations

import math

import numpy as np

from .integrands import DiagonalRidge, GaussianMixture


def gaussian_4d():
    """4-D Gaussian N exp(-100 |x - 0.5|^2) on [0,1]^4, integral 1
    (N = (100/pi)^2 / erf(5)^4); neval 1e5, nitn 10, FP64."""
    N = (100 / math.pi) ** 2 / math.erf(5) ** 4
    f = GaussianMixture([N], [100.0], [[0.5] * 4])
    return dict(name="4-D Gaussian", dim=4, f=f, neval=1e5, nitn=10, exact=1.0,
                precision="f64", ncomp=1)


def double_peak_8d():
    """Two equal-weight isotropic Gaussians, sigma 0.02, centred at 0.33*1 and
    0.67*1 in [0,1]^8, integral 1; neval 1e7/iteration, nitn 15."""
    f = GaussianMixture.isotropic([[0.33] * 8, [0.67] * 8], 0.02)
    return dict(name="8-D double peak", dim=8, f=f, neval=1e7, nitn=15, exact=1.0,
                precision="f32", ncomp=2)


def ridge_8d(diagonal=False):
    """Mean of 1000 isotropic Gaussians, sigma 0.005, centres evenly spaced on
    the diagonal from 0.25*1 to 0.75*1 in [0,1]^8; neval 2e7, nitn 20."""
    t = 0.25 + 0.5 * np.arange(1000) / 999.0
    sigma = 0.005
    A = np.full(1000, 1e-3 * (2 * math.pi * sigma ** 2) ** -4)
    B = np.full(1000, 1.0 / (2 * sigma ** 2))
    f = DiagonalRidge(A, B, t, 8) if diagonal else GaussianMixture(A, B, t[:, None] * np.ones(8))
    return dict(name="8-D ridge (1000 Gaussians)", dim=8, f=f, neval=2e7, nitn=20, exact=1.0,
                precision="f32", ncomp=1000, sigma=sigma)


def ridge_9d(diagonal=False):
    """The learned-map workload's integrand: mean of 500 diagonal Gaussians,
    sigma 0.01, centres on the diagonal from 0.25*1 to 0.75*1 in [0,1]^9."""
    t = 0.25 + 0.5 * np.arange(500) / 499.0
    sigma = 0.01
    A = np.full(500, (1 / 500) * (2 * math.pi * sigma ** 2) ** -4.5)
    B = np.full(500, 1.0 / (2 * sigma ** 2))
    f = DiagonalRidge(A, B, t, 9) if diagonal else GaussianMixture(A, B, t[:, None] * np.ones(9))
    return dict(name="9-D ridge (500 Gaussians)", dim=9, f=f, neval=2e7, nitn=20, exact=1.0,
                precision="f32", ncomp=500, sigma=sigma)


def flops_per_eval(dim, ncomp):
    """Algorithmic FP32 work per integrand evaluation:
    ncomp*(3 dim + 4) for the Gaussians + 9 dim + 9 for y, map, weight, sums."""
    return ncomp * (3 * dim + 4) + 9 * dim + 9


def batched_gaussians(n=4096, seed=12345):
    """The batched workload: n independent single-Gaussian integrals; dim ~ U{2..10},
    centre ~ U[0.2,0.8]^dim, sigma log-uniform in [0.01, 0.1], neval
    log-uniform in [1e4, 1e6]; exact value from the erf product."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        d = int(rng.integers(2, 11))
        c = rng.uniform(0.2, 0.8, d)
        s = math.exp(rng.uniform(math.log(0.01), math.log(0.1)))
        neval = int(round(math.exp(rng.uniform(math.log(1e4), math.log(1e6)))))
        exact = 1.0
        for cd in c:
            exact *= 0.5 * (math.erf((1 - cd) / (math.sqrt(2) * s)) + math.erf(cd / (math.sqrt(2) * s)))
        out.append(dict(dim=d, center=c, sigma=s, neval=neval, exact=exact))
    return out
