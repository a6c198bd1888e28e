"""GPU parity of AdaptiveMap kernels against the FP64 oracle.

Bar: interval indices and FP64 x/J bit-exact for identical y (no FMA
contraction in either build); training histograms bit-exact in ordered mode;
adapted grids within 1e-12 relative (exact mode within a few ulp: the device
log/pow differ from glibc's in the last bit).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def va():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import vegas_amd
    return vegas_amd


def _maps(va, orc, grid, ninc=None):
    return va.AdaptiveMap(grid, ninc=ninc), orc.OracleMap(grid, ninc=ninc)


def _edge_ys(rng, n, dim, ninc):
    y = rng.uniform(0, 1, (n, dim))
    y[0] = 0.0
    y[1] = 1.0
    # values one ulp either side of interval edges
    for d in range(dim):
        k = rng.integers(1, ninc[d], 40)
        e = k / ninc[d]
        y[2:42, d] = np.nextafter(e, 0.0)
        y[42:82, d] = np.nextafter(e, 1.0)
        y[82:122, d] = e
    return y


def test_map_fixture(va):
    # tests/test_vegas.py:76-95
    m = va.AdaptiveMap([[0, 1, 3], [-2, 0, 6]])
    y = np.array([[0, 0], [0.25, 0.25], [0.5, 0.5], [0.75, 0.75], [1.0, 1.0]])
    x = np.empty(y.shape)
    jac = np.empty(5)
    m.map(y, x, jac)
    np.testing.assert_allclose(x, [[0, -2], [0.5, -1], [1, 0], [2, 3], [3, 6]])
    np.testing.assert_allclose(jac, [8, 8, 48, 48, 48])
    np.testing.assert_allclose(m(y), x)
    np.testing.assert_allclose(m.jac(y), jac)


@pytest.mark.parametrize("ninc", [[7, 10, 5], [1000, 990, 996, 1000], [3]])
def test_map_f64_bitexact(va, orc, ninc):
    rng = np.random.default_rng(len(ninc))
    dim = len(ninc)
    grid = [np.concatenate([[lo], np.sort(rng.uniform(lo, lo + 2, 5)), [lo + 2]])
            for lo in rng.uniform(-3, 3, dim)]
    m, o = _maps(va, orc, grid, ninc=ninc)
    assert np.array_equal(m.grid[:, :max(ninc) + 1], o.grid[:, :max(ninc) + 1])
    y = _edge_ys(rng, 5000, dim, ninc)
    x, jac = m.map(y)
    xo, jo = o.map(y)
    assert np.array_equal(x.cpu().numpy(), xo)
    assert np.array_equal(jac.cpu().numpy(), jo)
    assert np.array_equal(m.jac1d(y), o.jac1d(y))


def test_map_f32_indices(va, orc):
    import torch
    rng = np.random.default_rng(7)
    ninc = [997, 1000, 13]
    grid = [[0, 0.3, 1], [-1, 5], [2, 2.5, 3, 8]]
    m, o = _maps(va, orc, grid, ninc=ninc)
    y32 = _edge_ys(rng, 20000, 3, ninc).astype(np.float32)
    x, jac = m.map(torch.as_tensor(y32).cuda())
    xo, jo = o.map(y32.astype(np.float64))
    # FP64 arithmetic on identical (float-valued) y: results are the oracle's
    # values rounded to float
    assert np.array_equal(x.cpu().numpy(), xo.astype(np.float32))
    assert np.array_equal(jac.cpu().numpy(), jo.astype(np.float32))


def test_invmap(va, orc):
    rng = np.random.default_rng(3)
    grid = [np.sort(rng.uniform(0, 1, 31)), np.sort(rng.uniform(-5, 2, 17))]
    m, o = _maps(va, orc, grid)
    x = np.concatenate([rng.uniform(-6, 3, (2000, 2)), o.map(rng.uniform(0, 1, (2000, 2)))[0]])
    y, jac = m.invmap(x)
    yo, jo = o.invmap(x)
    np.testing.assert_allclose(y.cpu().numpy(), yo, rtol=1e-15, atol=0)
    np.testing.assert_allclose(jac.cpu().numpy(), jo, rtol=1e-15)
    m2 = va.AdaptiveMap([[0., 1., 3.], [-2., 0., 6.]])  # tests/test_vegas.py:97-112
    yy, jj = m2.invmap(np.array([[0., -2.], [0.5, -1], [1, 0], [2, 3], [3, 6]]))
    np.testing.assert_allclose(yy.cpu().numpy(), [[0, 0], [.25, .25], [.5, .5], [.75, .75], [1, 1]])
    np.testing.assert_allclose(jj.cpu().numpy(), [8., 8, 48, 48, 48])


@pytest.mark.parametrize("n", [0, 1, 999, 70001])
def test_train_hist_ordered_bitexact(va, orc, n):
    rng = np.random.default_rng(n)
    ninc = [7, 1000, 64]
    m, o = _maps(va, orc, [[0, 1], [0, 2], [-1, 1]], ninc=ninc)
    y = rng.uniform(0, 1, (n, 3))
    if n > 10:
        y[:3] = 0.0
        y[3:6] = 1.0
        y[6:50, 1] = rng.integers(0, 1000, 44) / 1000.0
    f = rng.normal(size=n) * np.exp(rng.uniform(-20, 20, n))
    for _ in range(2):  # two calls: sums continue from the previous state
        m.add_training_data(y, f)
        o.add_training_data(y, f)
    if n == 0:
        assert np.array_equal(m.sum_f, np.zeros_like(m.sum_f))
        return
    for d, k in enumerate(ninc):
        assert np.array_equal(m.sum_f[d, :k], o.sum_f[d, :k])
        assert np.array_equal(m.n_f[d, :k], o.n_f[d, :k])


def test_train_hist_fast(va, orc):
    rng = np.random.default_rng(9)
    ninc = [50, 996, 7]
    m, o = _maps(va, orc, [[0, 1]] * 3, ninc=ninc)
    y = rng.uniform(0, 1, (300000, 3))
    f = rng.exponential(size=300000)
    m.add_training_data(y, f, ordered=False)
    o.add_training_data(y, f)
    for d, k in enumerate(ninc):
        np.testing.assert_allclose(m.sum_f[d, :k], o.sum_f[d, :k], rtol=1e-12)
        assert np.array_equal(m.n_f[d, :k], o.n_f[d, :k])


def _train(m, rng, F, n=4000):
    y = rng.uniform(0, 1, (n, m.dim))
    x, jac = m.map(y)
    x = x.cpu().numpy() if hasattr(x, 'cpu') else x
    jac = jac.cpu().numpy() if hasattr(jac, 'cpu') else jac
    return y, (F(x) * jac) ** 2


@pytest.mark.parametrize("alpha", [0.5, 1.5, -1.0, 0.0])
def test_adapt_exact_vs_oracle(va, orc, alpha):
    rng = np.random.default_rng(int(10 * abs(alpha)) + 1)
    ninc = [20, 996, 333]
    m, o = _maps(va, orc, [[0, 1], [-1, 2], [0, 5]], ninc=ninc)
    F = lambda x: np.exp(-30 * ((x - 0.4) ** 2).sum(1))  # noqa: E731
    for it in range(3):
        y, w = _train(o, rng, F)
        m.add_training_data(y, w)
        o.add_training_data(y, w)
        m.adapt(alpha=alpha)
        o.adapt(alpha=alpha)
        np.testing.assert_allclose(m.grid, o.grid, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("alpha", [0.5, 2.0])
def test_adapt_parallel_vs_oracle(va, orc, alpha):
    import torch
    rng = np.random.default_rng(4)
    ninc = [1000, 994, 10, 3]
    grid = [[0, 1], [-1, 2], [0, 5], [0, 1]]
    m, o = _maps(va, orc, grid, ninc=ninc)
    F = lambda x: np.exp(-50 * ((x[:, :2] - 0.3) ** 2).sum(1)) + 0.1  # noqa: E731
    y, w = _train(o, rng, F, 100000)
    m.add_training_data(y, w)
    o.add_training_data(y, w)
    m.adapt(alpha=alpha, exact=False)
    o.adapt(alpha=alpha)
    # 1e-12 of the axis span: a node near x = 0 carries the absolute error
    # of the span, not of its own magnitude
    for d, k in enumerate(ninc):
        span = o.grid[d, k] - o.grid[d, 0]
        np.testing.assert_allclose(m.grid[d, :k + 1], o.grid[d, :k + 1], rtol=0, atol=1e-12 * span)


def test_adapt_closed_forms(va):
    # tests/test_vegas.py:169-207
    g = 1. / 3. ** 0.5
    yg = [(1 - g) / 4., (1 + g) / 4, (3 - g) / 4, (3 + g) / 4.]
    m = va.AdaptiveMap([[0, 2]], ninc=2)
    y = np.array([[a] for a in yg])
    for _ in range(60):
        x, jac = m.map(y)
        m.add_training_data(y, x.cpu().numpy()[:, 0] ** 2 * jac.cpu().numpy())
        m.adapt(alpha=2.)
    np.testing.assert_allclose(m.grid, [[0, 2. / 2. ** (1. / 3.), 2.]])
    for alpha, nitn in [(2., 60), (-2., 20)]:
        m = va.AdaptiveMap([[0, 2], [0, 4]], ninc=2)
        y = np.array([[a, b] for a in yg for b in yg])
        for _ in range(nitn):
            x, jac = m.map(y)
            x = x.cpu().numpy()
            m.add_training_data(y, x[:, 0] * x[:, 1] ** 2 * jac.cpu().numpy())
            m.adapt(alpha=alpha, exact=(alpha < 0))
        np.testing.assert_allclose(m.grid, [[0, 2. * 2. ** (-0.5), 2.], [0, 4. * 2 ** (-1. / 3.), 4.]])


def test_adapt_ninc_and_quirk(va, orc):
    # regrid to new ninc preserves the map; alpha > 0 without training data
    # chains the shared buffer across axes (pyx:534)
    m, o = _maps(va, orc, [[0, 0.4, 1], [-2, 0., 4]], ninc=4)
    np.testing.assert_allclose(m.grid, [[0, 0.2, 0.4, 0.7, 1.], [-2., -1., 0., 2., 4.]])
    m, o = _maps(va, orc, [[0, 1], [0, 1], [0, 3]], ninc=[4, 6, 9])
    m.adapt(alpha=0.5)
    o.adapt(alpha=0.5)
    np.testing.assert_allclose(m.grid, o.grid, rtol=1e-13)
    m.adapt(ninc=[1, 1, 1])
    np.testing.assert_allclose(m.grid, [[0, 1], [0, 1], [0, 3]])


def _oracle_adapt_to_samples(o, x, fx, nitn, alpha, maxinc):
    """The reference's adapt_to_samples (pyx:772-804) spelled out on the
    oracle: regrid to tmp_ninc, nitn rounds of invmap -> add_training_data
    ((jac f)^2) -> adapt, then regrid to max(ninc, maxinc_axis)."""
    old_ninc = max(max(int(v) for v in o.ninc), maxinc)
    tmp_ninc = int(min(old_ninc, x.shape[0] / 10.))
    o.adapt(ninc=tmp_ninc)
    for _ in range(nitn):
        y, jac = o.invmap(x)
        o.add_training_data(y, (jac * fx) ** 2)
        o.adapt(alpha=alpha, ninc=tmp_ninc)
    if tmp_ninc != old_ninc:
        o.adapt(ninc=old_ninc)


@pytest.mark.parametrize("nx,alpha", [(5000, 1.0), (30000, 0.5)])
def test_adapt_to_samples_vs_oracle(va, orc, nx, alpha):
    """AdaptiveMap.adapt_to_samples (pyx:728-804) on the device equals the
    oracle's invmap / add_training_data / adapt chain: 1e-12 relative."""
    rng = np.random.default_rng(nx)
    grid = [[0, 1], [-1, 2], [0, 5]]
    x = np.stack([rng.beta(2, 5, nx), -1 + 3 * rng.beta(5, 2, nx), 5 * rng.uniform(0, 1, nx)], axis=1)
    f = lambda x: np.exp(-10 * (x[:, 0] - 0.3) ** 2) * (1 + x[:, 1] ** 2)  # noqa: E731
    m, o = _maps(va, orc, grid, ninc=[50, 60, 7])
    m.adapt_to_samples(x, f, nitn=4, alpha=alpha)
    _oracle_adapt_to_samples(o, x, f(x), 4, alpha, va.Integrator.defaults['maxinc_axis'])
    assert list(m.ninc) == [int(v) for v in o.ninc] == [1000] * 3
    np.testing.assert_allclose(m.grid, o.grid, rtol=1e-12, atol=1e-14)
    # f given as values, and the maxinc_axis default read from Integrator
    saved = va.Integrator.defaults['maxinc_axis']
    try:
        va.Integrator.defaults['maxinc_axis'] = 300
        m2, o2 = _maps(va, orc, grid, ninc=[50, 60, 7])
        m2.adapt_to_samples(x, f(x), nitn=2, alpha=alpha)
        _oracle_adapt_to_samples(o2, x, f(x), 2, alpha, 300)
        assert list(m2.ninc) == [300] * 3
        np.testing.assert_allclose(m2.grid, o2.grid, rtol=1e-12, atol=1e-14)
    finally:
        va.Integrator.defaults['maxinc_axis'] = saved
    with pytest.raises(ValueError):
        m.adapt_to_samples(x[:15], f)
