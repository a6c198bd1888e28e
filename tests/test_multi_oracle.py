"""The multi-component oracle (oracle/multi_oracle.py, pyx:2136-2183 with
fcn_size > 1) pinned by textbook identities, by the scalar rule of the C
oracle, and by the guard's closed form; and the host running average of
arrays (vegas_amd._ravg.RAvgArray, pyx:2453-2831) by the cases that reduce
to the scalar RAvg.  CPU only."""
import numpy as np
import pytest

from oracle import multi_oracle as mo


def _points(seed, nh=40, m=3):
    rng = np.random.default_rng(seed)
    n_h = rng.integers(2, 9, nh)
    hcube = np.repeat(np.arange(nh), n_h)
    w = rng.uniform(0.5, 1.5, len(hcube)) / len(hcube)
    f = rng.normal(size=(len(hcube), m)) + np.arange(1, m + 1)
    return hcube, w, f, n_h


def test_identities_away_from_guard():
    """Per hcube: var term = n * sample covariance of w f (ddof 1) and mean
    term = sum w f; summed over hcubes."""
    hcube, w, f, n_h = _points(1)
    mean, cov = mo.iteration_sums(hcube, w, f, correlate=True)
    wf = w[:, None] * f
    want_m = wf.sum(0)
    want_c = np.zeros((3, 3))
    a = 0
    for n in n_h:
        want_c += n * np.cov(wf[a:a + n].T, ddof=1)
        a += n
    np.testing.assert_allclose(mean, want_m, rtol=1e-13)
    np.testing.assert_allclose(cov, want_c, rtol=1e-11, atol=1e-300)
    m2, v2 = mo.iteration_sums(hcube, w, f, correlate=False)
    np.testing.assert_allclose(m2, mean, rtol=0, atol=0)
    np.testing.assert_allclose(v2, np.diag(cov), rtol=1e-13)


def test_scaled_component():
    """f1 = 3 f0: var1 = 9 var0, cov10 = 3 var0, correlation exactly 1."""
    hcube, w, f, _ = _points(2, m=1)
    f2 = np.concatenate([f, 3.0 * f], axis=1)
    mean, cov = mo.iteration_sums(hcube, w, f2)
    assert abs(mean[1] - 3 * mean[0]) <= 1e-13 * abs(mean[1])
    assert abs(cov[1, 1] - 9 * cov[0, 0]) <= 1e-12 * cov[1, 1]
    assert abs(cov[1, 0] - 3 * cov[0, 0]) <= 1e-12 * cov[1, 0]
    assert cov[0, 1] == cov[1, 0]


def test_constant_component_takes_the_guard():
    """A constant component with equal weights in each hcube: every
    deviation is guarded, var_h = n^2 (EPSILON |mean_h|)^2 / (n - 1), and
    it drops out of the covariances."""
    hcube, w, f, n_h = _points(3, m=2)
    wc = np.repeat(1.0 / (len(n_h) * n_h), n_h)
    f[:, 1] = 2.5
    mean, cov = mo.iteration_sums(hcube, wc, f)
    want = 0.0
    for n in n_h:
        mh = 2.5 / len(n_h) / n
        want += n * n * (mo.EPSILON * mh) ** 2 / (n - 1.)
    assert abs(cov[1, 1] - want) <= 1e-12 * want
    assert cov[1, 0] == 0.0
    assert abs(mean[1] - 2.5) <= 1e-14


def test_one_component_matches_c_oracle(orc):
    """m = 1 against the C oracle's scalar accumulation on the same points:
    beta = 0 (n_h = neval / nhcube) and a uniform map (w = 1/(nhcube n_h)),
    so the oracle's points, hcubes and weights are known to the test."""
    dim, nstrat, neval = 2, [5, 4], 2000
    o = orc.OracleIntegrator([(0, 1)] * dim, neval=neval, nstrat=nstrat, seed=31, beta=0.0, alpha=0.0)
    nh = int(np.prod(nstrat))
    n_h = neval // nh
    xs = []

    def fcn(x):
        xs.append(np.array(x))
        return np.exp(-8 * ((x - 0.4) ** 2).sum(1))

    out, _, _ = o.iterate(fcn)
    x = np.concatenate(xs)
    assert len(x) == nh * n_h == out.last_neval
    hcube = np.repeat(np.arange(nh), n_h)
    w = np.full(len(x), 1.0 / (nh * n_h))
    mean, var = mo.iteration_sums(hcube, w, fcn(x)[:, None], correlate=False)
    assert abs(mean[0] - out.mean) <= 1e-13 * abs(out.mean)
    assert abs(var[0] - out.var) <= 1e-11 * out.var


def test_ravg_array_reduces_to_scalar():
    """Diagonal covariances: the weighted array average equals the scalar
    weighted average per component (chi2 adds up, dof = n_itn*m - m);
    unweighted: mean of means, covariance sum/n^2."""
    from vegas_amd._ravg import RAvg, RAvgArray, GVar
    rng = np.random.default_rng(4)
    for weighted in (True, False):
        r = RAvgArray((3,), weighted=weighted)
        s = [RAvg(weighted) for _ in range(3)]
        for _ in range(5):
            m = rng.normal(size=3) + 5
            v = rng.uniform(0.01, 0.2, 3)
            r.add((m, np.diag(v)))
            for k in range(3):
                s[k].add(GVar(m[k], np.sqrt(v[k])))
        np.testing.assert_allclose(r.mean, [x.mean for x in s], rtol=1e-12)
        np.testing.assert_allclose(r.sdev, [x.sdev for x in s], rtol=1e-12)
        if weighted:
            assert abs(r.chi2 - sum(x.chi2 for x in s)) <= 1e-10 * r.chi2
            assert r.dof == 12


def test_ravg_array_correlated_weighted_closed_form():
    """Two iterations with the same covariance C: the weighted average is
    the plain mean with covariance C/2 (closed form of the GLS average)."""
    from vegas_amd._ravg import RAvgArray
    C = np.array([[0.04, 0.018], [0.018, 0.09]])
    r = RAvgArray((2,), weighted=True)
    r.add((np.array([1.0, 2.0]), C))
    r.add((np.array([1.2, 1.7]), C))
    np.testing.assert_allclose(r.mean, [1.1, 1.85], rtol=1e-12)
    np.testing.assert_allclose(r.cov, C / 2, rtol=1e-10)
    d = np.array([0.2, -0.3])
    assert abs(r.chi2 - 0.5 * d @ np.linalg.solve(C, d)) <= 1e-10


def test_ravg_array_pickle_and_summary():
    import pickle
    from vegas_amd._ravg import RAvgArray, RAvgDict
    rng = np.random.default_rng(6)
    r = RAvgArray((2,), weighted=True)
    for _ in range(3):
        m = rng.normal(size=2) + 3
        r.add((m, np.array([[0.04, 0.01], [0.01, 0.09]])))
    r.sum_neval = 300
    r2 = pickle.loads(pickle.dumps(r))
    np.testing.assert_allclose(r2.mean, r.mean, rtol=0, atol=0)
    np.testing.assert_allclose(r2.cov, r.cov, rtol=0, atol=0)
    assert r2.sum_neval == 300 and r2.nitn == 3
    text = r.summary(extended=True)
    assert text.splitlines()[0].split()[:2] == ['itn', 'integral'] and len(text.splitlines()) == 2 + 3 + 1 + 2
    d = RAvgDict([('a', (), 0, 1), ('b', (2,), 1, 3)])
    d.add((np.array([1.0, 2.0, 3.0]), np.diag([0.1, 0.2, 0.3])))
    d2 = pickle.loads(pickle.dumps(d))
    assert list(d2.keys()) == ['a', 'b'] and d2['b'].shape == (2,)
    assert abs(d2['a'].mean - 1.0) < 1e-15
