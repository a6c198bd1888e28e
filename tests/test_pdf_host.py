"""Host logic of PDFIntegrator and restratify (vegas_amd/pdf.py), on the
CPU: the Gaussian change of variables, first-order error propagation of the
ratio E = F / P, the batch conventions of user functions, and the strata
re-division.  Each is checked against what the mathematics fixes (the
covariance rebuilt from L, the determinant, a finite-difference Jacobian of
the ratio, products and orderings of the new strata)."""
import numpy as np
import pytest

pdf = pytest.importorskip("vegas_amd.pdf")


def test_gauss_pdf_change_of_variables():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((4, 4))
    cov = A @ A.T + 0.1 * np.eye(4)
    mean = rng.standard_normal(4)
    g = pdf._GaussPDF(mean, cov, svdcut=1e-15)
    np.testing.assert_allclose(g.L @ g.L.T, cov, rtol=1e-12, atol=1e-12)
    assert abs(g.dp_dchiv - np.sqrt(np.linalg.det(cov))) <= 1e-12 * g.dp_dchiv
    np.testing.assert_array_equal(g.pflat(np.zeros((1, 4)))[0], mean)
    # a unit chiv step moves p along a column of L
    e = np.zeros((1, 4))
    e[0, 2] = 1.0
    np.testing.assert_allclose(g.pflat(e)[0] - mean, g.L[:, 2], rtol=1e-14)


def test_gauss_pdf_svdcut_floor():
    cov = np.array([[1.0, 1.0], [1.0, 1.0]])  # singular
    g = pdf._GaussPDF([0.0, 0.0], cov, svdcut=1e-6)
    lam = np.linalg.eigvalsh(np.linalg.inv(np.diag([1.0, 1.0])) @ (g.L @ g.L.T))
    assert lam.min() > 0 and abs(lam.min() / lam.max() - 1e-6) < 1e-9


def test_ratio_propagation_matches_finite_differences():
    rng = np.random.default_rng(2)
    mean = np.array([2.0, 6.0, -3.0, 0.5])  # P at index 0
    B = rng.standard_normal((4, 4)) * 0.1
    cov = B @ B.T
    E, cE = pdf._ratio(mean, cov, 0)
    np.testing.assert_allclose(E, mean[1:] / mean[0])
    J = np.zeros((3, 4))
    h = 1e-6
    for k in range(4):
        dp = mean.copy()
        dm = mean.copy()
        dp[k] += h
        dm[k] -= h
        J[:, k] = (dp[1:] / dp[0] - dm[1:] / dm[0]) / (2 * h)
    np.testing.assert_allclose(cE, J @ cov @ J.T, rtol=1e-7, atol=1e-14)
    # P last (adapt_to_pdf=False layout)
    E2, cE2 = pdf._ratio(np.roll(mean, -1), np.roll(np.roll(cov, -1, 0), -1, 1), 3)
    np.testing.assert_allclose(E2, E)
    np.testing.assert_allclose(cE2, cE, rtol=1e-12)


def test_call_user_conventions():
    from vegas_amd.batchintegrand import lbatchintegrand, rbatchintegrand
    p = np.arange(12.0).reshape(4, 3)

    @lbatchintegrand
    def fl(q):
        return q[:, 0] * 2

    @rbatchintegrand
    def fr(q):
        return dict(a=q[0] * 2, b=q[1:])

    def fs(q):
        return q.sum()

    np.testing.assert_array_equal(pdf._call_user(fl, p), p[:, 0] * 2)
    r = pdf._call_user(fr, p)
    np.testing.assert_array_equal(r['a'], p[:, 0] * 2)
    np.testing.assert_array_equal(r['b'], p[:, 1:])
    np.testing.assert_array_equal(pdf._call_user(fs, p), p.sum(1))
    cols, (kind, lay) = pdf._flatten_f(r, 4)
    assert kind == 'dict' and cols.shape == (4, 3)
    assert [(k, s, a, b) for k, s, a, b in lay] == [('a', (), 0, 1), ('b', (2,), 1, 3)]


def test_restratified_strata():
    old = np.array([8, 8, 8, 8])
    # one dominant axis takes the strata; the product never grows
    new = pdf._restratified([1e6, 1.0, 1.0, 1.0], old, below_avg_nstrat=1)
    assert list(new[1:]) == [1, 1, 1] and new[0] <= np.prod(old) and new[0] > 0.9 * np.prod(old)
    # larger weight, more strata (gamma = 1); gamma = 0 keeps the old ones
    new = pdf._restratified([16.0, 4.0, 2.0, 1.0], old)
    assert new[0] >= new[1] >= new[2] >= new[3] >= 1 and np.prod(new) <= np.prod(old)
    new0 = pdf._restratified([16.0, 4.0, 2.0, 1.0], old, gamma=0.0)
    assert np.all(np.abs(new0 - old) <= 1)


def test_pdf_statistics_closed_forms_and_propagation():
    """PDFStatistics from raw moments: the standard normal (1, 0, 3 for
    E[z^2], E[z^3], E[z^4]) has skew 0 and excess kurtosis 0; z^2 for a
    standard normal z is chi-square with one degree of freedom, raw moments
    1, 3, 15, 105: mean 1, sdev sqrt(2), skew sqrt(8), excess kurtosis 12.
    The propagated errors equal J C J^T with J by central differences."""
    s = pdf.PDFStatistics([0.0, 1.0, 0.0, 3.0])
    assert (s.mean.mean, s.sdev.mean, s.skew.mean, s.ex_kurt.mean) == (0.0, 1.0, 0.0, 0.0)
    s = pdf.PDFStatistics([1.0, 3.0, 15.0, 105.0])
    np.testing.assert_allclose([s.mean.mean, s.sdev.mean, s.skew.mean, s.ex_kurt.mean],
                               [1.0, np.sqrt(2.0), np.sqrt(8.0), 12.0], rtol=1e-14)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((4, 4)) * 1e-3
    cov = A @ A.T
    m = np.array([0.3, 1.2, 0.9, 4.1])
    s = pdf.PDFStatistics(m, cov)

    def vals(mm):
        t = pdf.PDFStatistics(mm)
        return np.array([t.mean.mean, t.sdev.mean, t.skew.mean, t.ex_kurt.mean])
    h = 1e-6
    J = np.array([(vals(m + h * e) - vals(m - h * e)) / (2 * h) for e in np.eye(4)]).T
    want = np.sqrt(np.diag(J @ cov @ J.T))
    got = [s.mean.sdev, s.sdev.sdev, s.skew.sdev, s.ex_kurt.sdev]
    np.testing.assert_allclose(got, want, rtol=1e-6)
    assert 'ex_kurt' in str(s)


def test_hist_counts_and_median():
    """_hist_counts is one-hot per column with underflow first and overflow
    last (searchsorted side='right', __init__.py:1036-1042), so its column
    sums are numpy's histogram with the two outer bins added; the median
    read from bin probabilities of a uniform distribution on [0, 1] with
    edges 0, 0.25, ..., 1 is 0.5, and 0.3 after moving mass below 0."""
    rng = np.random.default_rng(3)
    v = rng.standard_normal((1000, 2))
    bins = np.array([np.linspace(-2, 2, 9), np.linspace(-1, 3, 9)])
    c = pdf._hist_counts(v, bins)
    assert c.shape == (1000, 2, 10) and np.all(c.sum(axis=2) == 1)
    for j in range(2):
        inner = np.histogram(v[:, j], bins[j])[0].astype(float)
        # numpy's last bin is closed; the right-side rule puts the edge value in overflow
        tot = c[:, j].sum(axis=0)
        assert tot[0] == np.sum(v[:, j] < bins[j][0]) and tot[-1] == np.sum(v[:, j] >= bins[j][-1])
        np.testing.assert_array_equal(tot[1:-1], inner)
    edges = np.linspace(0, 1, 5)
    s = pdf.PDFStatistics([0.5, 1 / 3], histogram=(edges, [0, .25, .25, .25, .25, 0], np.zeros(6)))
    assert abs(s.median - 0.5) < 1e-15 and s.skew is None and 'median' in str(s)
    s = pdf.PDFStatistics([0.5, 1 / 3], histogram=(edges, [0.2, .25, .25, .15, .15, 0], np.zeros(6)))
    assert abs(s.median - 0.3) < 1e-12
