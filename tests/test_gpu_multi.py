"""Multi-component integrands on the GPU (pyx:2136-2183 with fcn_size > 1,
correlate_integrals on and off; RAvgArray / RAvgDict results): the device
sums against the oracle on the same points, component 0 steering the
sampling exactly as a scalar integrand would, and analytic moments."""
import math

import numpy as np
import pytest

from oracle import multi_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def va():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import vegas_amd
    return vegas_amd


def _gauss(x):
    return np.exp(-30.0 * ((x - 0.45) ** 2).sum(1))


def _moments(va):
    @va.lbatchintegrand
    def f(x):
        g = _gauss(x)
        return np.stack([g, x[:, 0] * g, np.ones(len(x)), x[:, 1] * x[:, 1] * g], axis=1)
    return f


def _first_iteration_points(va, dim, neval, seed, **kw):
    twin = va.Integrator([[0, 1]] * dim, neval=neval, seed=seed, **kw)
    xs, ws, hs = [], [], []
    for x, w, h in twin._device_batches(yield_hcube=True):
        xs.append(x.cpu().numpy())
        ws.append(w.cpu().numpy())
        hs.append(h.cpu().numpy())
    return np.concatenate(xs), np.concatenate(ws), np.concatenate(hs)


@pytest.mark.parametrize("correlate", [True, False])
def test_device_sums_match_oracle(va, correlate):
    dim, neval, seed = 3, 20000, 13
    f = _moments(va)
    integ = va.Integrator([[0, 1]] * dim, neval=neval, seed=seed, correlate_integrals=correlate)
    r = integ(f, nitn=1)
    x, w, h = _first_iteration_points(va, dim, neval, seed)
    assert len(x) == integ.last_neval
    mean, cov = mo.iteration_sums(h, w, f(x), correlate=correlate)
    g = r.itn_results[0]
    np.testing.assert_allclose(g.mean, mean, rtol=1e-12, atol=1e-15)
    if correlate:
        np.testing.assert_allclose(g.cov, cov, rtol=1e-10, atol=1e-300)
    else:
        np.testing.assert_allclose(np.diag(g.cov), cov, rtol=1e-10, atol=1e-300)
        assert np.count_nonzero(g.cov - np.diag(np.diag(g.cov))) == 0


@pytest.mark.parametrize("m", [4, 6])
@pytest.mark.parametrize("correlate", [True, False])
def test_device_sums_mixed_hcube_sizes(va, m, correlate):
    """Hcubes from 2 to several hundred points in one batch (a lognormal
    sigf steers the allocation): short hcubes summed per thread, long ones
    by the whole workgroup; m = 4 takes the all-pairs kernel, m = 6 the
    per-pair kernel."""
    dim, neval, seed = 3, 40000, 29
    rng = np.random.default_rng(5)
    kw = dict(nstrat=[8, 8, 4], neval_frac=0.95, sigf=np.exp(2.5 * rng.standard_normal(256)))

    @va.lbatchintegrand
    def f(x):
        g = _gauss(x)
        cols = [g, x[:, 0] * g, np.ones(len(x)), x[:, 1] * x[:, 1] * g, x[:, 2] * g, 1.0 + x[:, 0]]
        return np.stack(cols[:m], axis=1)

    integ = va.Integrator([[0, 1]] * dim, neval=neval, seed=seed, correlate_integrals=correlate, **kw)
    r = integ(f, nitn=1, adapt=False)
    x, w, h = _first_iteration_points(va, dim, neval, seed, **kw)
    n_h = np.bincount(h, minlength=256)
    assert n_h.min() <= 32 < n_h.max()  # both paths taken
    mean, cov = mo.iteration_sums(h, w, f(x), correlate=correlate)
    g = r.itn_results[0]
    np.testing.assert_allclose(g.mean, mean, rtol=1e-12, atol=1e-15)
    if correlate:
        np.testing.assert_allclose(g.cov, cov, rtol=1e-10, atol=1e-300)
    else:
        np.testing.assert_allclose(np.diag(g.cov), cov, rtol=1e-10, atol=1e-300)


def test_component_zero_steers_sampling(va):
    """The multi-component run draws the same points, sigf and grid as the
    scalar run of its component 0: after three iterations the component-0
    estimates agree to summation order."""
    dim, neval = 3, 30000

    @va.lbatchintegrand
    def f0(x):
        return _gauss(x)

    a = va.Integrator([[0, 1]] * dim, neval=neval, seed=5)
    b = va.Integrator([[0, 1]] * dim, neval=neval, seed=5)
    ra = a(f0, nitn=3)
    rb = b(_moments(va), nitn=3)
    for k in range(3):
        ga, gb = ra.itn_results[k], rb.itn_results[k]
        assert abs(ga.mean - gb.mean[0]) <= 1e-12 * abs(ga.mean)
        assert abs(ga.sdev - gb.sdev[0]) <= 1e-9 * ga.sdev
    np.testing.assert_allclose(a.map.grid, b.map.grid, rtol=0, atol=1e-13)


def test_dict_integrand_and_moments(va):
    """Dictionary output (RAvgDict with the integrand's keys and shapes) and
    analytic values: int g, int x0 g and int x1^2 g of the Gaussian
    g = exp(-30 |x - 0.45|^2) on [0,1]^3, and the volume 1 (within 5 sigma)."""
    dim = 3
    a = 30.0

    def one(c):  # int_0^1 exp(-a (x - c)^2) dx
        return 0.5 * math.sqrt(math.pi / a) * (math.erf(math.sqrt(a) * (1 - c)) + math.erf(math.sqrt(a) * c))

    def first(c):  # int_0^1 x exp(-a (x - c)^2) dx
        return c * one(c) + (math.exp(-a * c * c) - math.exp(-a * (1 - c) ** 2)) / (2 * a)

    def second(c):  # int_0^1 x^2 exp(-a (x - c)^2) dx
        t = lambda x: -(x + c) * math.exp(-a * (x - c) ** 2) / (2 * a)
        return (t(1) - t(0)) + (c * c + 1 / (2 * a)) * one(c)

    I0 = one(0.45) ** 3
    exact = {'g': I0, 'mom': [first(0.45) * one(0.45) ** 2, second(0.45) * one(0.45) ** 2], 'vol': 1.0}

    @va.lbatchintegrand
    def f(x):
        g = _gauss(x)
        return {'g': g, 'mom': np.stack([x[:, 0] * g, x[:, 1] ** 2 * g], axis=1), 'vol': np.ones(len(x))}

    integ = va.Integrator([[0, 1]] * dim, neval=40000, seed=8)
    integ(f, nitn=5)
    r = integ(f, nitn=8)
    assert list(r.keys()) == ['g', 'mom', 'vol']
    assert r['mom'].shape == (2,)
    assert abs(r['g'].mean - exact['g']) < 5 * r['g'].sdev
    for k in range(2):
        assert abs(r['mom'].mean[k] - exact['mom'][k]) < 5 * r['mom'].sdev[k]
    # sum w estimates the volume; with an adapted map it fluctuates with J
    assert abs(r['vol'].mean - 1.0) < 5 * r['vol'].sdev
    assert r.dof > 0 and 0.0 <= r.Q <= 1.0


def test_array_shape_preserved(va):
    """(n, 2, 2) outputs give a (2, 2) RAvgArray; (n, 1) a (1,) one."""
    @va.lbatchintegrand
    def f(x):
        g = _gauss(x)
        return np.stack([g, 2 * g, 3 * g, 4 * g], axis=1).reshape(-1, 2, 2)

    r = va.Integrator([[0, 1]] * 2, neval=5000, seed=3)(f, nitn=3)
    assert r.mean.shape == (2, 2)
    np.testing.assert_allclose(r.mean.reshape(-1) / r.mean.reshape(-1)[0], [1, 2, 3, 4], rtol=1e-10)

    @va.lbatchintegrand
    def f1(x):
        return _gauss(x)[:, None]

    r1 = va.Integrator([[0, 1]] * 2, neval=5000, seed=3)(f1, nitn=3)
    assert r1.mean.shape == (1,)
