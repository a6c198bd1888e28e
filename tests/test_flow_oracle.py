"""The FP64 coupling-flow oracle pinned by closed forms and finite
differences (CPU only): inverse o forward = identity, log J = log|det| of
the numerical Jacobian, and backprop() = finite differences of the loss."""
import numpy as np
import pytest

from oracle import flow_oracle as fo


def _params(dim, seed, scale=0.3):
    P = fo.init_params(dim, 6, seed)
    rng = np.random.default_rng(seed + 1)
    for p in P:  # a non-trivial map: nonzero output layer and biases
        p['W3'] = rng.normal(0, scale / 10, p['W3'].shape)
        p['b3'] = rng.normal(0, scale, p['b3'].shape)
        p['b2'] = rng.normal(0, 0.1, p['b2'].shape)
        p['W1'][:, 15] = rng.normal(0, 0.1, 128)
    return P, fo.masks(dim, 6)


def test_masks_split_every_pair():
    for dim in (2, 3, 5, 9, 15):
        M = fo.masks(dim, 6)
        for a in M:
            assert a.any() or dim == 1
        for i in range(dim):
            for j in range(i + 1, dim):
                assert any(m[i] != m[j] for m in M), (dim, i, j)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.14159, 0.0])
    r = fo.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0            # tie -> even
    assert r[2] == 1.0 + 2 ** -7 + 2 ** -8 - 2 ** -8 or abs(r[2] - (1 + 2 ** -7)) < 1e-12
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert r[4] == 0.0


@pytest.mark.parametrize("bf", [False, True])
def test_inverse_undoes_forward(bf):
    P, M = _params(9, 3)
    z = np.random.default_rng(0).uniform(0.05, 0.95, (200, 9))
    x, lj = fo.forward(P, M, z, bf)
    z2, lq = fo.inverse(P, M, x, bf)
    # with bf16 GEMM inputs a last-bit change in a conditioning value can
    # flip its bf16 rounding between the two directions
    tol = 1e-6 if bf else 1e-8
    np.testing.assert_allclose(z2, z, rtol=0, atol=tol)
    np.testing.assert_allclose(lq, -lj, rtol=tol, atol=tol)


def test_identity_at_init():
    P = fo.init_params(5, 6, 0)
    M = fo.masks(5, 6)
    z = np.random.default_rng(1).uniform(0.01, 0.99, (50, 5))
    x, lj = fo.forward(P, M, z, False)
    np.testing.assert_allclose(x, z, atol=1e-12)
    np.testing.assert_allclose(lj, 0.0, atol=1e-12)


def test_logdet_matches_numerical_jacobian():
    dim = 4
    P, M = _params(dim, 7, scale=0.2)
    rng = np.random.default_rng(2)
    z = rng.uniform(0.1, 0.9, (5, dim))
    x, lj = fo.forward(P, M, z, False)
    h = 1e-6
    for n in range(len(z)):
        J = np.zeros((dim, dim))
        for d in range(dim):
            zp, zm = z[n:n + 1].copy(), z[n:n + 1].copy()
            zp[0, d] += h
            zm[0, d] -= h
            J[:, d] = (fo.forward(P, M, zp, False)[0][0] - fo.forward(P, M, zm, False)[0][0]) / (2 * h)
        assert abs(np.log(abs(np.linalg.det(J))) - lj[n]) < 1e-6


def test_backprop_matches_finite_differences():
    dim = 5
    P, M = _params(dim, 11, scale=0.5)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.05, 0.95, (40, dim))
    c = rng.uniform(0.0, 1.0, 40)
    c /= c.sum()
    grads, _ = fo.backprop(P, M, x, c, bf=False)
    h = 1e-6
    for k in (0, 2, 5):
        for name, idx in (('W1', (3, 1)), ('W1', (7, 15)), ('W2', (5, 9)), ('b2', (17,)),
                          ('W3', (1, 4)), ('W3', (17, 8)), ('b3', (2,)), ('b3', (16,))):
            p = P[k][name]
            old = p[idx]
            p[idx] = old + h
            lp = fo.loss(P, M, x, c)
            p[idx] = old - h
            lm = fo.loss(P, M, x, c)
            p[idx] = old
            fd = (lp - lm) / (2 * h)
            g = grads[k][name][idx]
            assert abs(g - fd) <= 1e-6 * max(1.0, abs(fd)), (k, name, idx, g, fd)
