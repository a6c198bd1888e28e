"""GPU vegas+ iterations against the FP64 oracle and analytic values.

FP64 parity mode (precision='f64', exact=True) feeds the device and the
oracle the same Philox uniforms, so the first iteration's y, x, J and weights
are identical; the estimate, variance and sigf agree to 1e-12 relative
(the integrand's exp differs in the last bit), and the training histogram is
bit for bit the oracle's when the integrand is evaluated identically.  Later
iterations are compared statistically: a one-ulp sigf difference may flip an
int() in the allocation and re-deal the random numbers.
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


def gauss4(va, orc):
    N = (100 / math.pi) ** 2 / math.erf(5) ** 4
    return va.GaussianMixture([N], [100.], [[0.5] * 4]), orc.GaussMix([N], [100.], [[0.5] * 4])


def double_peak(va, dim=8, sigma=0.02):
    return va.GaussianMixture.isotropic([[0.33] * dim, [0.67] * dim], sigma)


@pytest.mark.parametrize("neval,nstrat", [(20000, None), (3300, [5, 5, 4, 4]), (2e5, None)])
def test_first_iteration_parity(va, orc, neval, nstrat):
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=neval, nstrat=nstrat, seed=11, precision='f64',
                          exact=True, alpha=0.0)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=neval, nstrat=nstrat, seed=11, alpha=0.0)
    out, sf, nf = o.iterate(fo)
    assert integ.last_neval == out.last_neval
    m = r.itn_results[0]
    assert abs(m.mean - out.mean) <= 1e-12 * abs(out.mean)
    assert abs(m.var - out.var) <= 1e-10 * out.var
    np.testing.assert_allclose(integ.sigf, o.sigf, rtol=1e-11)
    assert abs(integ.sum_sigf - o.sigf.sum()) <= 1e-10 * o.sigf.sum()


DIMS = [1, 2, 3, 5, 7, 9, 12, 13, 17]


@pytest.mark.parametrize("dim", DIMS)
def test_first_iteration_parity_dims(va, orc, dim):
    """Every dimension instantiation of the fused sampler (1-12 unrolled,
    13+ the runtime-dim path): FP64 parity mode against the oracle."""
    c = [0.5 + 0.02 * ((d % 3) - 1) for d in range(dim)]
    f = va.GaussianMixture([1.0], [8.0], [c])
    fo = orc.GaussMix([1.0], [8.0], [c])
    integ = va.Integrator([[0, 1]] * dim, neval=20000, seed=23, precision='f64', exact=True, alpha=0.0)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * dim, neval=20000, seed=23, alpha=0.0)
    out, sf, nf = o.iterate(fo)
    assert integ.last_neval == out.last_neval
    m = r.itn_results[0]
    assert abs(m.mean - out.mean) <= 1e-12 * abs(out.mean)
    assert abs(m.var - out.var) <= 1e-10 * out.var
    np.testing.assert_allclose(integ.sigf, o.sigf, rtol=1e-11)


@pytest.mark.parametrize("dim", DIMS)
def test_fused_equals_staged_f32_dims(va, dim):
    """24-bit uniforms (the integer stratum path wherever the axes allow):
    the fused sampler with its FP32 functor and the staged path with a
    numpy integrand see the same points, so their first-iteration estimates
    agree to the FP32 integrand's rounding."""
    c = [0.5 + 0.02 * ((d % 3) - 1) for d in range(dim)]
    f = va.GaussianMixture([1.0], [8.0], [c])
    cc = np.array(c)

    @va.lbatchintegrand
    def fl(x):
        return np.exp(-8.0 * ((x - cc) ** 2).sum(1))

    ra = va.Integrator([[0, 1]] * dim, neval=20000, seed=29, precision='f32')(f, nitn=1)
    rb = va.Integrator([[0, 1]] * dim, neval=20000, seed=29, precision='f32')(fl, nitn=1)
    for a, b in zip(ra.itn_results, rb.itn_results):
        assert abs(a.mean - b.mean) <= 2e-5 * abs(b.mean)
        assert abs(a.sdev - b.sdev) <= 1e-3 * b.sdev


@pytest.mark.parametrize("wl", ["gaussian_4d", "double_peak_8d"])
def test_fp32_path_error_against_oracle(va, orc, wl):
    """The FP32 path (24-bit uniforms, FP32 integrand functor; map, weights
    and sums in FP64) against the FP64 oracle fed the same 24-bit uniforms
    (rng_f32): the same points, so the difference is the FP32 integrand's
    rounding (and the map's integer stratum path rounding y once less).
    Bars: estimate 2e-5 relative, sdev 1e-3 relative, identical allocation;
    the measured errors are printed."""
    from vegas_amd import workloads
    w = getattr(workloads, wl)()
    f, dim = w["f"], w["dim"]
    fo = orc.GaussMix(f.A, f.B, f.centers)
    neval = 2e5
    integ = va.Integrator([[0, 1]] * dim, neval=neval, seed=19, precision='f32', alpha=0.0)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * dim, neval=neval, seed=19, alpha=0.0, rng_f32=True)
    out, _, _ = o.iterate(fo)
    assert integ.last_neval == out.last_neval
    m = r.itn_results[0]
    em = abs(m.mean - out.mean) / abs(out.mean)
    es = abs(m.sdev - math.sqrt(out.var)) / math.sqrt(out.var)
    print("FP32 path vs FP64 oracle (%s, neval %d): estimate %.2e relative, sdev %.2e relative"
          % (wl, integ.last_neval, em, es))
    assert em <= 2e-5 and es <= 1e-3, (em, es)


def test_fp32_path_far_tails_against_oracle(va, orc):
    """The FP32 Gaussian functor keeps the FP64 exponent range: on a narrow
    10-D peak (sigma 0.015) sampled by 2e4 uniform points, the first
    iteration's estimate is made of far-tail values (the nearest points
    near e^-500, below anything FP32 represents, so FP32 exp would give 0)
    and still equals the FP64 oracle's on the same 24-bit points to the
    FP32 rounding of the exponent (about 1e-7 of it: ~1e-4 relative here)."""
    dim = 10
    f = va.GaussianMixture.isotropic([[0.5] * dim], 0.015)
    fo = orc.GaussMix(f.A, f.B, f.centers)
    integ = va.Integrator([[0, 1]] * dim, neval=2e4, seed=29, precision='f32', alpha=0.0)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * dim, neval=2e4, seed=29, alpha=0.0, rng_f32=True)
    out, _, _ = o.iterate(fo)
    assert integ.last_neval == out.last_neval
    m = r.itn_results[0]
    assert 0 < out.mean < 1e-46  # the estimate is all tails, below FP32's range
    em = abs(m.mean - out.mean) / out.mean
    print("far tails, FP32 path vs FP64 oracle: %.3g vs %.3g, %.2e relative" % (m.mean, out.mean, em))
    assert em <= 1e-3, em


@pytest.mark.parametrize("neval,scale", [(30000, 3.0), (2e5, 6.0)])
def test_parity_variable_nh(va, orc, neval, scale):
    """Heavy-tailed sigf: n_h from 2 to thousands, hcubes spanning many wave
    tiles and cut at every lane offset."""
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=neval, seed=17, precision='f64', exact=True,
                          alpha=0.0)
    rng = np.random.default_rng(int(neval))
    sigf = np.exp(scale * rng.standard_normal(integ.nhcube))
    integ.set(sigf=sigf)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=neval, seed=17, alpha=0.0)
    o.sigf = sigf.copy()
    o.sum_sigf = float(np.sum(sigf))
    out, sf, nf = o.iterate(fo)
    assert integ.last_neval == out.last_neval
    assert integ.neval_hcube_range[1] == out.nh_max and out.nh_max > 600
    m = r.itn_results[0]
    assert abs(m.mean - out.mean) <= 1e-12 * abs(out.mean)
    assert abs(m.var - out.var) <= 1e-10 * out.var
    np.testing.assert_allclose(integ.sigf, o.sigf, rtol=1e-11)


def test_adapt_to_errors_parity(va, orc):
    """adapt_to_errors=True (pyx:1688-1692, 2186-2193): uniform allocation,
    the grid trained on each hcube's variance; first iteration and the
    adapted grid against the oracle in FP64 parity mode."""
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=30000, seed=31, precision='f64', exact=True,
                          adapt_to_errors=True)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=30000, seed=31, adapt_to_errors=True)
    m, v, nev = o.run(fo, 1)
    assert integ.last_neval == nev[0]
    g = r.itn_results[0]
    assert abs(g.mean - m[0]) <= 1e-12 * abs(m[0])
    assert abs(g.var - v[0]) <= 1e-10 * v[0]
    np.testing.assert_allclose(integ.map.grid, o.map.grid, rtol=1e-11, atol=1e-13)


def test_no_adaptive_stratification_parity(va, orc):
    """beta = 0 (pyx:1663-1666): every hcube gets the same count and sigf is
    not used; the map still adapts.  First iteration and the adapted grid."""
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=30000, seed=37, precision='f64', exact=True, beta=0.0)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=30000, seed=37, beta=0.0)
    m, v, nev = o.run(fo, 1)
    assert integ.last_neval == nev[0]
    g = r.itn_results[0]
    assert abs(g.mean - m[0]) <= 1e-12 * abs(m[0])
    assert abs(g.var - v[0]) <= 1e-10 * v[0]
    np.testing.assert_allclose(integ.map.grid, o.map.grid, rtol=1e-11, atol=1e-13)


def test_max_neval_hcube_clamp_parity(va, orc):
    """max_neval_hcube caps n_h (pyx:1696-1698): heavy-tailed sigf against
    a cap of 40 points per hcube."""
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=30000, seed=41, precision='f64', exact=True, alpha=0.0,
                          max_neval_hcube=40)
    rng = np.random.default_rng(9)
    sigf = np.exp(3.0 * rng.standard_normal(integ.nhcube))
    integ.set(sigf=sigf)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=30000, seed=41, alpha=0.0, max_neval_hcube=40)
    o.sigf = sigf.copy()
    o.sum_sigf = float(np.sum(sigf))
    out, sf, nf = o.iterate(fo)
    assert integ.last_neval == out.last_neval
    assert integ.neval_hcube_range[1] == out.nh_max == 40
    g = r.itn_results[0]
    assert abs(g.mean - out.mean) <= 1e-12 * abs(out.mean)
    assert abs(g.var - out.var) <= 1e-10 * out.var


def test_adapt_false_iterations_and_unweighted_average(va, orc):
    """adapt=False: nothing is trained, every iteration samples the same
    grid and allocation with a fresh stream, and RAvg is the unweighted
    average (pyx:2392-2410); all three iterations match the oracle's."""
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=20000, seed=43, precision='f64', exact=True, adapt=False)
    r = integ(f, nitn=3)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=20000, seed=43, adapt=False)
    m, v, nev = o.run(fo, 3)
    for k in range(3):
        assert abs(r.itn_results[k].mean - m[k]) <= 1e-12 * abs(m[k])
        assert abs(r.itn_results[k].var - v[k]) <= 1e-10 * v[k]
    mean, sdev, chi2 = orc.ravg(m, v, weighted=False)
    assert abs(r.mean - mean) <= 1e-12 * abs(mean)
    assert abs(r.sdev - sdev) <= 1e-10 * sdev
    assert abs(r.chi2 - chi2) <= 1e-8 * max(chi2, 1e-300)


def test_training_and_adapt_parity(va, orc):
    # one iteration with adaptation: histogram -> adapt on device vs oracle
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=50000, seed=5, precision='f64', exact=True)
    integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=50000, seed=5)
    o.run(fo, 1)
    np.testing.assert_allclose(integ.map.grid, o.map.grid, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("neval,nstrat", [(20000, None), (3300, [5, 5, 4, 4]), (2e5, None)])
def test_training_histogram_exact_bitexact(va, orc, neval, nstrat):
    """FP64 parity mode: the training histogram of an iteration (sum_f of
    (w f n)^2 and n_f per increment, pyx:421-464 fed by pyx:2173-2196) is
    the oracle's bit for bit: the sampler's records are summed per bin in
    point order, as the reference's sequential loop is.  The integrands are
    evaluated identically on both sides (the device exp may differ from
    glibc's in the last bit): a numpy polynomial through the staged path,
    and a constant device integrand on a non-uniform grid, where the
    Jacobian makes every (w f n)^2 different."""
    @va.lbatchintegrand
    def poly(x):
        return x[:, 0] * x[:, 1] * x[:, 1] + 0.5 * x[:, 2] * x[:, 3] + 0.25

    def poly_o(x):
        return x[:, 0] * x[:, 1] * x[:, 1] + 0.5 * x[:, 2] * x[:, 3] + 0.25

    ninc = orc.geometry(4, neval=neval, nstrat=nstrat)["ninc"]
    nodes = [list(np.linspace(0, 1, int(n) + 1) ** (1.0 + 0.3 * d)) for d, n in enumerate(ninc)]
    uniform = [list(np.linspace(0, 1, int(n) + 1)) for n in ninc]
    cases = [(poly, poly_o, uniform),
             (va.GaussianMixture([1.0], [0.0], [[0.5] * 4]), orc.GaussMix([1.0], [0.0], [[0.5] * 4]), nodes)]
    for f, fo, grid in cases:
        integ = va.Integrator(va.AdaptiveMap(grid), neval=neval, nstrat=nstrat, seed=11, precision='f64',
                              exact=True, alpha=0.5)
        integ(f, nitn=1)
        o = orc.OracleIntegrator([(0, 1)] * 4, neval=neval, nstrat=nstrat, seed=11, alpha=0.5)
        o.map = orc.OracleMap(grid)
        assert list(o.map.ninc) == list(ninc)
        out, sf, nf = o.iterate(fo)
        assert integ.last_neval == out.last_neval
        gs, gn = integ._last_training()
        assert gs.shape == sf.shape and np.count_nonzero(sf) > 0.9 * sf.size
        assert np.array_equal(gn, nf)
        assert np.array_equal(gs, sf), np.max(np.abs(gs - sf) / np.maximum(sf, 1e-300))


def test_training_histogram_fast_mode(va, orc):
    """Default mode (53-bit uniforms for an FP64 integrand): the histogram
    built from the records in a fixed order has the oracle's counts exactly
    and its sums to 1e-12, on the same points (y differs from the oracle's
    division by at most an ulp, which moves no point across a bin edge
    here: the counts are equal)."""
    f, fo = gauss4(va, orc)
    integ = va.Integrator([[0, 1]] * 4, neval=2e5, seed=13, precision='f64', alpha=0.5)
    integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * 4, neval=2e5, seed=13, alpha=0.5)
    _, sf, nf = o.iterate(fo)
    gs, gn = integ._last_training()
    assert np.array_equal(gn, nf)
    np.testing.assert_allclose(gs, sf, rtol=1e-12, atol=0)


def test_unordered_histogram_option(va):
    """deterministic=False (LDS atomics in timing order) trains on the same
    points: the same first-iteration estimate, the same counts, sums to
    rounding, and an estimate consistent with the default mode's."""
    f = double_peak(va, dim=8, sigma=0.05)
    a = va.Integrator([[0, 1]] * 8, neval=4e5, seed=77)
    b = va.Integrator([[0, 1]] * 8, neval=4e5, seed=77, deterministic=False)
    ra, rb = a(f, nitn=1), b(f, nitn=1)
    assert ra.mean == rb.mean and ra.sdev == rb.sdev
    (sa, na), (sb, nb) = a._last_training(), b._last_training()
    assert np.array_equal(na, nb)
    np.testing.assert_allclose(sb, sa, rtol=1e-12, atol=0)
    ra, rb = a(f, nitn=3), b(f, nitn=3)
    assert abs(ra.mean - rb.mean) < 5 * math.hypot(ra.sdev, rb.sdev)


def test_training_histogram_large_batch(va):
    """One batch of 2e7 points in 8-D (FP32 integrand, integer stratum path):
    each workgroup takes more than 224 tiles, so the u16 LDS counts are moved
    into its u32 row while the tiles are still computed.  The default
    (tile-ordered) histogram is bit for bit the same on a second run with the
    same seed; the unordered one has the same counts and sums to rounding;
    every axis' counts add up to the points sampled (less the rare y = 0
    exactly, which train no bin)."""
    f = double_peak(va, dim=8, sigma=0.05)
    a = va.Integrator([[0, 1]] * 8, neval=2e7, seed=31)
    b = va.Integrator([[0, 1]] * 8, neval=2e7, seed=31)
    c = va.Integrator([[0, 1]] * 8, neval=2e7, seed=31, deterministic=False)
    ra, rb, rc = a(f, nitn=1), b(f, nitn=1), c(f, nitn=1)
    assert ra.mean == rb.mean == rc.mean and ra.sdev == rb.sdev == rc.sdev
    (sa, na), (sb, nb), (sc, nc) = a._last_training(), b._last_training(), c._last_training()
    assert np.array_equal(sa, sb) and np.array_equal(na, nb)
    assert np.array_equal(na, nc)
    np.testing.assert_allclose(sc, sa, rtol=1e-12, atol=0)
    npts = a.last_neval
    assert npts > 224 * 256 * 256
    tot = na.sum(axis=1)
    assert np.all(tot <= npts) and np.all(tot >= npts - 200), (tot, npts)


def test_y_zero_left_out_of_training(va, orc):
    """A point with y = 0 exactly (one stratum per axis, 24-bit uniform 0 on
    axis 0: seed 15, point 477932) stays out of axis 0's training histogram
    (pyx:459-463) on the integer stratum-to-bin path.  Constant integrand,
    non-uniform grid on axis 1: the adapted grid equals the oracle's to
    rounding, far below the shift that counting the point would cause."""
    seed, p, neval = 15, 477932, 600000
    w = orc.philox([p, 0, 0, 0], [seed, 0])
    assert w[0] >> 8 == 0
    grid = [[0.0, 1.0], [0.0, 0.05, 0.2, 1.0]]
    o = orc.OracleIntegrator(grid, neval=neval, nstrat=[1, 1], seed=seed, rng_f32=True)
    m0 = orc.OracleMap(o.map)
    out, sf, nf = o.iterate(orc.GaussMix([1.0], [0.0], [[0.5, 0.5]]))

    def adapted(s, n):
        m = orc.OracleMap(m0)
        m.sum_f, m.n_f = s.copy(), n.copy()
        m.adapt(alpha=0.5)
        return m.grid

    g_ok = adapted(sf, nf)
    # the same histogram with point p counted on axis 0 (w f n = J there)
    _, jp = m0.map(np.array([[0.0, (int(w[1]) >> 8) / 16777216.0]]))
    sf2, nf2 = sf.copy(), nf.copy()
    sf2[0, 0] += jp[0] ** 2
    nf2[0, 0] += 1
    gap = np.abs(adapted(sf2, nf2) - g_ok).max()
    assert gap > 1e-9
    f = va.GaussianMixture([1.0], [0.0], [[0.5, 0.5]])
    integ = va.Integrator(va.AdaptiveMap(grid), neval=neval, nstrat=[1, 1], seed=seed, precision='f32')
    integ(f, nitn=1)
    assert integ.last_neval == neval
    assert np.abs(integ.map.grid - g_ok).max() < gap / 100


def test_zero_uniform_in_upper_stratum_trained(va, orc):
    """A 24-bit uniform of 0 in stratum 1 of 2 gives y = 1/2, inside (0,1),
    so the point is trained on (pyx:459-463 leaves out y = 0 only): the
    sampler's screen for y = 0, which looks at the uniform first, must let it
    through.  Seed 63, point 121083 (uniform 0 on axis 0) with nstrat [2, 1]:
    the first iteration puts the same number of points in both hcubes, so the
    point is in hcube 1, stratum 1 on axis 0, and every point of that hcube
    lands in the upper half of axis 0's bins.  The device's training counts
    equal the oracle's bit for bit (the oracle draws the same 24-bit
    uniforms), and a point counted out would show as one short there."""
    seed, p, neval = 63, 121083, 600000
    assert orc.philox([p, 0, 0, 0], [seed, 0])[0] >> 8 == 0
    region = [[0.0, 1.0], [0.0, 1.0]]
    o = orc.OracleIntegrator(region, neval=neval, nstrat=[2, 1], seed=seed, rng_f32=True)
    out, sf, nf = o.iterate(orc.GaussMix([1.0], [0.0], [[0.5, 0.5]]))
    f = va.GaussianMixture([1.0], [0.0], [[0.5, 0.5]])
    integ = va.Integrator(region, neval=neval, nstrat=[2, 1], seed=seed, precision='f32')
    assert list(integ.map.ninc) == list(o.map.ninc) and integ.map.ninc[0] % 2 == 0
    integ(f, nitn=1)
    npts = out.last_neval
    assert integ.last_neval == npts and npts // 2 <= p < npts
    gs, gn = integ._last_training()
    half = int(o.map.ninc[0]) // 2
    assert round(nf[0, half:].sum()) == npts // 2
    assert np.array_equal(gn, nf)
    np.testing.assert_allclose(gs, sf, rtol=1e-12, atol=0)


def test_fused_equals_staged(va, orc):
    """Same seed and precision: the device integrand and a numpy lbatch,
    rbatch, torch and non-batch integrand see the same points."""
    import torch
    f, _ = gauss4(va, orc)
    N, B = f.A[0], f.B[0]

    @va.lbatchintegrand
    def fl(x):
        return N * np.exp(-B * ((x - 0.5) ** 2).sum(1))

    @va.rbatchintegrand
    def fr(x):
        return N * np.exp(-B * ((x - 0.5) ** 2).sum(0))

    @va.gpu_lbatchintegrand
    def ft(x):
        return N * torch.exp(-B * ((x - 0.5) ** 2).sum(1))

    def fs(x):
        return N * math.exp(-B * sum((xi - 0.5) ** 2 for xi in x))

    res = []
    for fn, kw in [(f, {}), (fl, dict(min_neval_batch=3000)), (fr, {}), (ft, {}), (fs, {})]:
        neval = 4000 if fn is fs else 30000
        integ = va.Integrator([[0, 1]] * 4, neval=neval, seed=3, precision='f64', **kw)
        res.append((neval, integ(fn, nitn=3)))
    ref = [r for n, r in res if n == 30000][0]
    for n, r in res:
        if n != 30000:
            continue
        for a, b in zip(r.itn_results, ref.itn_results):
            assert abs(a.mean - b.mean) <= 1e-12 * abs(b.mean)


def test_constant_integrand_guard(va):
    # tests/test_vegas.py:1272-1296
    @va.batchintegrand
    def g(x):
        return 7 * np.ones(x.shape[0], float)
    integ = va.Integrator([(0, 1)], neval=100, alpha=0)
    res = integ(g, nitn=4)
    assert round(res.itn_results[0].sdev / res.itn_results[1].sdev - 1.0, 7) == 0
    assert round(res.itn_results[0].sdev / res.sdev - 2.0, 7) == 0
    assert round(res.mean - 7.0, 7) == 0
    res = integ(g, nitn=4, neval=1e2, adapt=False)
    assert round(res.itn_results[0].sdev / res.itn_results[1].sdev - 1.0, 7) == 0
    assert round(res.itn_results[0].sdev / res.sdev - 2.0, 7) == 0

    @va.batchintegrand
    def z(x):
        return np.zeros(x.shape[0], float)
    assert va.Integrator([(0, 1)])(z, nitn=4, neval=100).mean == 0.0


def test_volume(va):
    # tests/test_vegas.py:774-781 with a device integrand (constant via B=0)
    f = va.GaussianMixture([2.0], [0.0], [[0.0, 0.0]])
    r = va.Integrator([[-1, 1], [0, 4]], seed=1)(f)
    assert abs(r.mean - 16) < 1e-6 * 16 and r.sdev < 1e-6


def test_nan_raises(va):
    @va.batchintegrand
    def f(x):
        v = np.ones(x.shape[0])
        v[min(3, len(v) - 1)] = np.nan  # the one-point shape probe included
        return v
    with pytest.raises(ValueError):
        va.Integrator([[0, 1]] * 2, neval=1000)(f, nitn=2)

    @va.batchintegrand
    def g(x):
        return 1 / 0.
    with pytest.raises(ZeroDivisionError):
        va.Integrator([[0, 1]] * 2, neval=1000)(g, nitn=2)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_gaussian_4d_statistics(va, prec):
    # 4-D Gaussian workload, neval 1e5, nitn 10
    N = (100 / math.pi) ** 2 / math.erf(5) ** 4
    f = va.GaussianMixture([N], [100.], [[0.5] * 4])
    integ = va.Integrator([[0, 1]] * 4, neval=1e5, seed=2026, precision=prec)
    integ(f, nitn=10)
    r = integ(f, nitn=10)
    assert abs(r.mean - 1.0) < 5 * r.sdev
    assert r.Q > 1e-3
    assert r.sdev < 2e-4


def test_double_peak_8d_statistics(va):
    # 8-D double-peak workload, neval 1e7, nitn 15 (FP32 integrand)
    f = double_peak(va)
    integ = va.Integrator([[0, 1]] * 8, neval=1e7, seed=7)
    assert list(integ.nstrat) == [6] * 6 + [5] * 2
    r = integ(f, nitn=15)
    rr = integ(f, nitn=5)
    assert abs(rr.mean - 1.0) < 5 * rr.sdev
    assert rr.Q > 1e-3
    assert abs(r.mean - 1.0) < 5 * max(r.sdev, 1e-3)
    # device running average agrees with the host one
    dev = integ._last_device_ravg
    assert abs(dev[0] - rr.mean) <= 1e-12 * abs(rr.mean)
    assert abs(dev[1] - rr.sdev) <= 1e-9 * rr.sdev


def test_random_batch_integral(va):
    # tests/test_vegas.py:1070-1087
    f = va.GaussianMixture([1.0], [100.], [[0.5] * 4])
    integ = va.Integrator(4 * [[0, 1]], seed=9)
    integ(f, nitn=10, neval=1000)
    result = integ(f, nitn=1, neval=1000, adapt=False)
    integral = 0.0
    for x, wgt in integ.random_batch():
        integral += wgt.dot(f(x))
    assert abs(result.mean - integral) < 5 * result.sdev


@pytest.mark.parametrize("dim,precision", [(4, None), (8, 'f32'), (5, 'f64')])
def test_seed_determinism(va, dim, precision):
    """Same seed, same bits: every iteration's estimate, the adapted grid,
    sigf and the last training histogram of two runs are identical (the
    training histogram is summed in a fixed order, independent of timing)."""
    f = double_peak(va, dim=dim, sigma=0.05)
    runs = []
    for _ in range(2):
        integ = va.Integrator([[0, 1]] * dim, neval=2e5, seed=42, precision=precision)
        r = integ(f, nitn=4)
        runs.append((r, integ))
    (a, ia), (b, ib) = runs
    assert len(a.itn_results) == 4
    for x, y in zip(a.itn_results, b.itn_results):
        assert x.mean == y.mean and x.sdev == y.sdev
    assert np.array_equal(ia.map.grid, ib.map.grid)
    assert np.array_equal(ia.sigf, ib.sigf)
    for u, v in zip(ia._last_training(), ib._last_training()):
        assert np.array_equal(u, v)


def test_copy_samples_independently(va):
    """Integrator(integ) is trained like integ but draws its own points (a
    fresh seed), as copies sharing the reference's global RNG do; seed= fixes
    the copy's stream."""
    f = double_peak(va, dim=4, sigma=0.05)
    a = va.Integrator([[0, 1]] * 4, neval=2e4, seed=42)
    a(f, nitn=2)
    b = va.Integrator(a)
    assert b.seed != a.seed
    np.testing.assert_array_equal(b.map.extract_grid()[0], a.map.extract_grid()[0])
    ra, rb = a(f, nitn=1), b(f, nitn=1)
    assert ra.itn_results[0].mean != rb.itn_results[0].mean
    c, d = va.Integrator(a, seed=7), va.Integrator(a, seed=7)
    assert c(f, nitn=1).mean == d(f, nitn=1).mean


def test_diagonal_ridge_first_iteration_parity(va, orc):
    """The diagonal-ridge functor (|x - t 1|^2 = P + d (xbar - t)^2) against
    the oracle's plain Gaussian mixture with centres t 1: FP64, same uniforms,
    so only the integrand's algebra and exp rounding differ."""
    dim, nc, sigma = 6, 20, 0.05
    t = 0.3 + 0.4 * np.arange(nc) / (nc - 1)
    A = np.full(nc, (2 * math.pi * sigma ** 2) ** (-dim / 2) / nc)
    B = np.full(nc, 1 / (2 * sigma ** 2))
    f = va.DiagonalRidge(A, B, t, dim)
    fo = orc.GaussMix(A, B, t[:, None] * np.ones(dim))
    integ = va.Integrator([[0, 1]] * dim, neval=2e5, seed=21, precision='f64', exact=True, alpha=0.0)
    r = integ(f, nitn=1)
    o = orc.OracleIntegrator([(0, 1)] * dim, neval=2e5, seed=21, alpha=0.0)
    out, sf, nf = o.iterate(fo)
    assert integ.last_neval == out.last_neval
    m = r.itn_results[0]
    assert abs(m.mean - out.mean) <= 1e-10 * abs(out.mean)
    assert abs(m.var - out.var) <= 1e-8 * out.var


def test_ridge_8d_workload(va):
    """The 8-D ridge workload (1000 Gaussians, sigma 0.005, neval 2e7) has
    the reference rule's geometry; the diagonal and generic functors give
    the same first iteration on the same points (FP32 rounding) on the same
    ridge widened to sigma 0.05.  (At sigma 0.005 the first iteration's
    estimate, about 2e-31, is made of far-tail values: the generic functor
    keeps them with the FP64 exponent range, the diagonal one flushes those
    below e^-104 as plain FP32 does, and the two differ there by about 1%.)"""
    from vegas_amd import workloads
    for diag in (True, False):
        wl = workloads.ridge_8d(diagonal=diag)
        integ = va.Integrator([[0, 1]] * 8, neval=wl["neval"], seed=2026, precision='f32')
        assert list(integ.nstrat) == [7, 7, 6, 6, 6, 6, 6, 6]
        assert integ.nhcube == 2286144
        assert list(integ.map.ninc) == [994, 994] + [996] * 6
    ncomp, sigma, dim = 1000, 0.05, 8
    t = 0.25 + 0.5 * np.arange(ncomp) / (ncomp - 1)
    A = np.full(ncomp, (2 * math.pi * sigma ** 2) ** (-dim / 2) / ncomp)
    B = np.full(ncomp, 1 / (2 * sigma ** 2))
    fd = va.DiagonalRidge(A, B, t, dim)
    res = {}
    for name, f in (('diag', fd), ('generic', fd.as_mixture())):
        integ = va.Integrator([[0, 1]] * dim, neval=2e6, seed=2026, precision='f32')
        r = integ(f, nitn=1)
        res[name] = r.itn_results[0]
        assert math.isfinite(res[name].mean) and res[name].sdev > 0
    assert abs(res['diag'].mean - res['generic'].mean) <= 1e-4 * abs(res['generic'].mean)
    assert abs(res['diag'].sdev - res['generic'].sdev) <= 1e-3 * res['generic'].sdev
