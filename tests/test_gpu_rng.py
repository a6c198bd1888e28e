"""The RNG seam: Integrator(ran_array_generator=...) (pyx:1081-1086,
1676-1680).  The reference calls ran_array_generator((neval_batch, dim)) once
per batch and forms y = (stratum + u)/nstrat from its rows in point order
(pyx:1732-1741); the staged device iteration reads the caller's numbers from
a per-iteration buffer in place of its Philox stream."""
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


class PhiloxReplay:
    """A generator that hands out the device's own 53-bit Philox uniforms
    (oracle.uniforms, whose bit layout is pinned against independent numpy
    code) for one iteration, batch after batch."""

    def __init__(self, orc, seed, dim):
        self.orc, self.seed, self.dim = orc, seed, dim
        self.start(0)

    def start(self, iteration):
        self.it, self.p = iteration, 0
        self.calls = []

    def __call__(self, size):
        n, d = size
        assert d == self.dim
        self.calls.append(n)
        u = self.orc.uniforms(self.seed, self.it, self.p, n, d)
        self.p += n
        return u


def _double_peak(dim):
    @__import__('vegas_amd').lbatchintegrand
    def f(x):
        r1 = np.sum((x - 0.33) ** 2, axis=1)
        r2 = np.sum((x - 0.67) ** 2, axis=1)
        return np.exp(-r1 / 0.005) + np.exp(-r2 / 0.005)
    return f


@pytest.mark.parametrize("exact", [False, True])
def test_replayed_stream_equals_default(va, orc, exact):
    """A generator that replays the Philox stream gives, bit for bit, the
    estimates, the adapted grid and sigf of the default generator over four
    iterations (two batches per iteration): the caller's row p is used
    exactly where the device would draw point p."""
    dim, seed = 3, 91
    f = _double_peak(dim)
    kw = dict(neval=6000, seed=seed, precision='f64', exact=exact, min_neval_batch=2500)
    a = va.Integrator([[0, 1]] * dim, **kw)
    ra = a(f, nitn=4)
    gen = PhiloxReplay(orc, seed, dim)
    b = va.Integrator([[0, 1]] * dim, ran_array_generator=gen, **kw)
    res = []
    for it in range(4):
        gen.start(it)
        res.append(b(f, nitn=1).itn_results[0])
        assert len(gen.calls) >= 2 and sum(gen.calls) == b.last_neval
    for x, y in zip(ra.itn_results, res):
        assert x.mean == y.mean and x.sdev == y.sdev
    assert np.array_equal(a.map.grid, b.map.grid)
    assert np.array_equal(a.sigf, b.sigf)


def test_random_batch_follows_reference_formula(va, orc):
    """random_batch with a numpy generator: y = (y0 + u)/nstrat with y0 the
    hcube's digits, axis 0 fastest (pyx:1735-1741), x and the jacobian from
    the map (pyx:1742), weight jac * dv_y / n_h (pyx:1746-1752) -- all
    bit for bit in exact mode, with the generator called once per batch with
    the batch's shape (pyx:1732)."""
    dim = 3
    calls = []
    rng = np.random.default_rng(2024)

    def gen(size):
        calls.append(size)
        return rng.random(size)

    integ = va.Integrator([[0, 1], [-1, 2], [0, 5]], neval=4000, seed=5, precision='f64', exact=True,
                          min_neval_batch=1500, ran_array_generator=gen)
    f = _double_peak(dim)
    integ(f, nitn=2)   # a trained, non-uniform grid and sigf
    calls.clear()
    # the replay starts where the two training iterations left the stream
    replay = np.random.default_rng(0)
    rng_state = rng.bit_generator.state
    batches = list(integ.random_batch(yield_hcube=True, yield_y=True))
    assert len(batches) >= 2 and all(c[1] == dim for c in calls)
    assert [c[0] for c in calls] == [b[0].shape[0] for b in batches]
    replay.bit_generator.state = rng_state
    m = integ.map
    om = orc.OracleMap([m.grid[d, :int(m.ninc[d]) + 1] for d in range(dim)])
    ns = np.asarray(integ.nstrat)
    hc_all = np.concatenate([b[3] for b in batches])
    n_h = np.bincount(hc_all, minlength=integ.nhcube)
    for x, y, w, hc in batches:
        u = replay.random((x.shape[0], dim))
        y0 = np.empty((x.shape[0], dim))
        t = hc.copy()
        for d in range(dim):
            y0[:, d] = t % ns[d]
            t = t // ns[d]
        y_ref = (y0 + u) / ns[None, :]
        assert np.array_equal(y, y_ref)
        x_ref, jac = om.map(y_ref)
        assert np.array_equal(x, x_ref)
        assert np.array_equal(w, jac * ((1.0 / integ.nhcube) / n_h[hc]))


def test_device_integrand_through_generator(va):
    """A device integrand with a generator runs the staged iteration through
    its batch form; the same numbers through a Python integrand give the
    same bits, and a pseudo-random stream integrates the normalised peak
    to 1 within 5 sigma."""
    dim = 4
    f = va.GaussianMixture.isotropic([[0.4] * dim], 0.1)

    @va.lbatchintegrand
    def g(x):
        return f(x)

    out = []
    for fcn in (f, g):
        integ = va.Integrator([[0, 1]] * dim, neval=20000, seed=3,
                              ran_array_generator=np.random.default_rng(77).random)
        out.append(integ(fcn, nitn=4))
    a, b = out
    for x, y in zip(a.itn_results, b.itn_results):
        assert x.mean == y.mean and x.sdev == y.sdev
    assert abs(a.mean - 1.0) < 5 * a.sdev


def test_generator_errors(va):
    """Wrong shapes and the 24-bit path are refused with ValueError; the
    default generator is restored with ran_array_generator=None."""
    f = _double_peak(2)
    integ = va.Integrator([[0, 1]] * 2, neval=1000, seed=1,
                          ran_array_generator=lambda size: np.zeros((size[0], size[1] + 1)))
    with pytest.raises(ValueError, match='ran_array_generator'):
        integ(f, nitn=1)
    integ.set(ran_array_generator=np.random.default_rng(1).random, precision='f32')
    with pytest.raises(ValueError, match='ran_array_generator'):
        integ(f, nitn=1)
    integ.set(ran_array_generator=None)
    r = integ(f, nitn=1)
    assert np.isfinite(r.mean)
