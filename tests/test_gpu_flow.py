"""The learned map on the GPU: the MFMA lane layout, the flow's forward
(sampling), inverse (log q) and gradient kernels against the FP64 oracle
(oracle/flow_oracle.py, with the same bf16 rounding points), and what a
trained flow gains over the vegas grid on a non-separable integrand."""
import math

import numpy as np
import pytest

from oracle import flow_oracle as fo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _params(dim, seed, scale=0.3):
    P = fo.init_params(dim, 6, seed)
    rng = np.random.default_rng(seed + 1)
    for p in P:
        p['W3'] = rng.normal(0, scale / 10, p['W3'].shape)
        p['b3'] = rng.normal(0, scale, p['b3'].shape)
        p['b2'] = rng.normal(0, 0.1, p['b2'].shape)
        p['W1'][:, 15] = rng.normal(0, 0.1, 128)
        p['W1'][:, dim:15] = 0.0
    return P


def _bf16_bits(a):
    b = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return (((b + 0x7FFF + ((b >> 16) & 1)) >> 16)).astype(np.uint16)


def test_mfma_lane_layout(torch):
    from vegas_amd import _lib
    from vegas_amd.flow import _bind
    L = _bind(_lib.lib())
    rng = np.random.default_rng(0)
    A = fo.bf16(rng.normal(size=(16, 16)))
    B = fo.bf16(rng.normal(size=(16, 16)))
    Ad = torch.as_tensor(_bf16_bits(A).astype(np.int16), device='cuda')
    Bd = torch.as_tensor(_bf16_bits(B).astype(np.int16), device='cuda')
    D = torch.zeros((16, 16), dtype=torch.float32, device='cuda')
    _lib.check(L.vgx_flow_mfma_probe(_lib.ptr(Ad), _lib.ptr(Bd), _lib.ptr(D), _lib.stream_ptr()), 'probe')
    np.testing.assert_allclose(D.cpu().numpy(), A @ B, rtol=1e-5, atol=1e-5)


def test_masks_match_oracle():
    from vegas_amd.flow import masks
    for dim in (2, 5, 9, 15):
        for a, b in zip(masks(dim, 6), fo.masks(dim, 6)):
            assert list(a) == list(b)


def test_logq_vs_oracle(torch):
    from vegas_amd.flow import CouplingFlow, pack
    dim = 9
    P = _params(dim, 3)
    fl = CouplingFlow(dim, 6, params=pack(P))
    x = np.random.default_rng(4).uniform(0.05, 0.95, (5003, dim))
    lq, u0 = fl.logq(x)
    _, lq_o, u0_o, _ = fo.inverse(P, fo.masks(dim, 6), x.astype(np.float32).astype(np.float64), True, keep=True)
    d = np.abs(lq.cpu().numpy() - lq_o)
    assert np.percentile(d, 90) < 1e-3, np.percentile(d, [50, 90, 99])
    assert d.max() < 5e-2
    du = np.abs(u0.cpu().numpy() - u0_o)
    assert np.percentile(du, 90) < 1e-3


def test_forward_vs_oracle(torch):
    """flow samples: x = F(z) and f J against the oracle's forward map on the
    same Philox z (counter (i, iteration, lane group), 24-bit, offset 2^-25)."""
    from oracle import oracle as orc
    import vegas_amd as va
    from vegas_amd.flow import CouplingFlow, pack
    dim, n, seed, itn = 9, 2048, 77, 3
    P = _params(dim, 5, scale=0.1)
    fl = CouplingFlow(dim, 6, params=pack(P))
    f = va.GaussianMixture([1.0], [2.0], [[0.5] * dim])
    s1, s2, xo, fj = fl.sample(f, n, seed=seed, iteration=itn, keep_points=True)
    z = np.zeros((n, dim))
    for i in range(n):
        for g in range((dim + 3) // 4):
            w = orc.philox([i, 0, itn, g], [seed & 0xffffffff, seed >> 32])
            for j in range(4):
                d = 4 * g + j
                if d < dim:
                    z[i, d] = (int(w[j]) >> 8) / 16777216.0 + 1.0 / 33554432.0
    x_o, lj_o = fo.forward(P, fo.masks(dim, 6), z, True)
    x_g = xo.cpu().numpy()[:, :dim]
    dx = np.abs(x_g - x_o)
    assert np.percentile(dx, 90) < 1e-4, np.percentile(dx, [50, 90, 99])
    fj_o = f(x_o) * np.exp(lj_o)
    assert np.all(np.isfinite(fj_o)) and np.all(fj_o > 1e-30)
    r = np.abs(fj.cpu().numpy() / fj_o - 1)
    assert np.percentile(r, 90) < 1e-3, np.percentile(r, [50, 90, 99])
    assert abs(s1 - fj.cpu().numpy().astype(np.float64).sum()) < 1e-4 * abs(s1)


def test_grad_vs_oracle(torch):
    from vegas_amd.flow import CouplingFlow, pack, unpack
    dim, n = 9, 4100  # not a multiple of the 16-sample tile
    P = _params(dim, 9, scale=0.1)
    fl = CouplingFlow(dim, 6, params=pack(P))
    rng = np.random.default_rng(10)
    x = rng.uniform(0.05, 0.95, (n, dim)).astype(np.float32).astype(np.float64)
    f = rng.uniform(0.1, 2.0, n).astype(np.float32).astype(np.float64)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32).astype(np.float64) / n
    M = fo.masks(dim, 6)
    _, lq = fo.inverse(P, M, x, True)
    lc = np.log(w) + 2 * np.log(f) - lq
    c = np.exp(lc - lc.max())
    c /= c.sum()
    g_o, _ = fo.backprop(P, M, x, c, bf=True)
    fl.grad(fl.padded(x), f, np.log(w))
    g = unpack(fl.last_grad())
    for k in range(6):
        for name in ('W1', 'W2', 'b2', 'W3', 'b3'):
            a, b = g[k][name], g_o[k][name]
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            assert err < 2e-2, (k, name, err)


@pytest.mark.parametrize("n", [1, 37, 300, 70000])
def test_grad_sizes_and_reproducible(torch, n):
    """The backward kernel forms the weight gradients in LDS, eight tiles of
    16 samples per workgroup round: a batch smaller than one tile, one that
    leaves most waves of the round without a tile, and one of several rounds
    per workgroup all give the oracle's gradients, and a second call gives
    the same bits (tiles go to workgroups in a fixed assignment)."""
    from vegas_amd.flow import CouplingFlow, pack, unpack
    dim = 7
    P = _params(dim, 3, scale=0.1)
    fl = CouplingFlow(dim, 6, params=pack(P))
    rng = np.random.default_rng(n)
    x = rng.uniform(0.05, 0.95, (n, dim)).astype(np.float32).astype(np.float64)
    f = rng.uniform(0.1, 2.0, n).astype(np.float32).astype(np.float64)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32).astype(np.float64) / n
    M = fo.masks(dim, 6)
    _, lq = fo.inverse(P, M, x, True)
    lc = np.log(w) + 2 * np.log(f) - lq
    c = np.exp(lc - lc.max())
    c /= c.sum()
    g_o, _ = fo.backprop(P, M, x, c, bf=True)
    xd = fl.padded(x)
    fl.grad(xd, f, np.log(w))
    G1 = fl.last_grad().copy()
    fl.grad(xd, f, np.log(w))
    assert np.array_equal(G1.view(np.uint32), fl.last_grad().view(np.uint32))
    g = unpack(G1)
    for k in range(6):
        for name in ('W1', 'W2', 'b2', 'W3', 'b3'):
            a, b = g[k][name], g_o[k][name]
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            assert err < 2e-2, (n, k, name, err)


def _ridge(dim, ncomp=50, sigma=0.03):
    import vegas_amd as va
    t = 0.3 + 0.4 * np.arange(ncomp) / (ncomp - 1)
    A = np.full(ncomp, (2 * math.pi * sigma ** 2) ** (-dim / 2) / ncomp)
    B = np.full(ncomp, 1 / (2 * sigma ** 2))
    return va.DiagonalRidge(A, B, t, dim)


def test_training_beats_vegas_grid(torch):
    """5-D diagonal ridge (non-separable): a flow trained on vegas samples
    (forward-KL recipe, about 500 Adam steps) has a per-sample variance well
    below the adapted vegas grid's, is accepted, and its estimate is
    consistent with the exact value 1 -- what the network gains over the
    grid alone."""
    import vegas_amd as va
    from vegas_amd.flow import FlowMap, vegas_training_data
    dim = 5
    f = _ridge(dim)
    integ = va.Integrator([[0, 1]] * dim, neval=1000000, seed=3)
    integ(f, nitn=10)
    r = integ(f, nitn=1)
    vegas_var = r.itn_results[0].sdev ** 2 * integ.last_neval
    fm = FlowMap(dim, seed=1)
    x, fx, lw = vegas_training_data(integ, f, fm.flow, 1000000)
    fm.train(f, x, fx, lw, epochs=50, batch=100000, lr=1e-3, clip=1.0, loss='kl')
    mean, var = fm.per_sample_variance(f, 1000000)
    assert fm.accept(f, vegas_var)
    assert var < 0.25 * vegas_var, (var, vegas_var)
    est, n = fm.integrate(f, rtol=1e-3, batch=1000000)
    assert abs(est.mean - 1.0) < 5 * est.sdev, (est.mean, est.sdev)


def test_grid_kept_when_flow_is_worse(torch):
    """The vegas grid is the fallback: on a separable Gaussian the untrained
    flow (the identity map, plain uniform sampling) loses and is rejected."""
    import vegas_amd as va
    from vegas_amd.flow import FlowMap
    dim = 4
    f = va.GaussianMixture([(100 / math.pi) ** 2], [100.0], [[0.5] * dim])
    integ = va.Integrator([[0, 1]] * dim, neval=100000, seed=2)
    integ(f, nitn=8)
    r = integ(f, nitn=1)
    vegas_var = r.itn_results[0].sdev ** 2 * integ.last_neval
    fm = FlowMap(dim, seed=1)
    assert not fm.accept(f, vegas_var)


def test_config5_ridge_9d_flow_beats_grid(torch):
    """The learned-map workload itself (9-D, 500 diagonal Gaussians, sigma
    0.01, integral 1) with the benchmark's recipe: the grid adapted over 10
    iterations of neval 2e7, the flow trained on 2e7 of its samples (3 KL
    epochs, 200 + 200 self-training steps).  The flow is accepted (lower
    per-sample variance than the adapted grid's own estimate of its
    variance) and its estimate is consistent with 1 at 1e-3."""
    import vegas_amd as va
    from vegas_amd import workloads
    from vegas_amd.flow import FlowMap
    wl = workloads.ridge_9d(diagonal=True)
    f = wl["f"]
    integ = va.Integrator([[0, 1]] * 9, neval=wl["neval"], seed=5)
    integ(f, nitn=10)
    r = integ(f, nitn=1)
    vegas_var = r.itn_results[0].sdev ** 2 * integ.last_neval
    fm = FlowMap(9, seed=1)
    fm.fit(f, integ, ntrain=2e7)
    assert fm.accept(f, vegas_var), (fm.flow_var_per_sample, vegas_var)
    est, n = fm.integrate(f, rtol=1e-3, batch=4000000, max_samples=4e9)
    print("9-D ridge: flow variance per sample %.3g, grid %.3g; flow estimate %.6f +- %.2g from %d samples"
          % (fm.flow_var_per_sample, vegas_var, est.mean, est.sdev, n))
    assert est.sdev < 1e-3 * abs(est.mean) and abs(est.mean - 1.0) < 5 * est.sdev, (est.mean, est.sdev, n)


def test_config3_ridge_8d_flow_and_grid(torch):
    """Config 3's ridge (8-D, 1000 diagonal Gaussians, sigma 0.005,
    integral 1): the adapted grid stays far from the integral (plain vegas+
    does not resolve it: below 0.5 after 20 iterations of neval 2e7), while
    the flow trained on the grid's samples (the benchmark's recipe with the
    adaptive variance phase) is accepted and its estimate is consistent with
    1 at 1e-3."""
    import vegas_amd as va
    from vegas_amd import workloads
    from vegas_amd.flow import FlowMap
    wl = workloads.ridge_8d(diagonal=True)
    f = wl["f"]
    integ = va.Integrator([[0, 1]] * 8, neval=wl["neval"], seed=5)
    integ(f, nitn=10)
    r = integ(f, nitn=10)
    assert r.itn_results[-1].mean < 0.5, r.itn_results[-1]
    vegas_var = r.itn_results[-1].sdev ** 2 * integ.last_neval
    fm = FlowMap(8, seed=1)
    fm.fit(f, integ, ntrain=2e7, rtol=1e-3)
    assert fm.accept(f, vegas_var), (fm.flow_var_per_sample, vegas_var)
    est, n = fm.integrate(f, rtol=1e-3, batch=4000000, max_samples=4e9)
    print("8-D ridge: grid's last iteration %.4f; flow variance per sample %.3g, grid %.3g; flow estimate "
          "%.6f +- %.2g from %d samples" % (r.itn_results[-1].mean, fm.flow_var_per_sample, vegas_var,
                                            est.mean, est.sdev, n))
    assert est.sdev < 1e-3 * abs(est.mean) and abs(est.mean - 1.0) < 5 * est.sdev, (est.mean, est.sdev, n)
