"""FP64 numpy restatement of the coupling-flow map (test infrastructure only).

The flow has no counterpart in the reference (vegas' map is the separable
grid, src/vegas/_vegas.pyx:39-832); this oracle defines the arithmetic the
HIP kernels in vegas_amd/csrc/vgx_flow.hip implement, so that their outputs
and gradients can be checked.  Only tests may import it.

Map.  z in (0,1)^D uniform -> u0 = logit(z) -> L affine couplings in R^D ->
x = sigmoid(uL).  Coupling k with conditioning set A_k (mask m_k) and
transformed set B_k:
    a   = [u_d if d in A_k else 0 for d < D] padded to 16, a[15] = 1 (bias)
    h1  = relu(W1 a)                  W1: [128][16]  (column 15 = bias)
    h2  = relu(W2 h1 + b2)            W2: [128][128]
    o   = W3 h2 + b3                  W3: [32][128]; rows 0..D-1 = raw s, rows 16..16+D-1 = t
    s_d = 4 tanh(o_d / 4),  t_d = o_{16+d}           (d in B_k; 0 on A_k)
    forward  u_B <- u_B exp(s) + t,       log J += sum_B s
    inverse  u_B <- (u_B - t) exp(-s),    log q -= sum_B s
log J(z) = sum_d [log sig'(uL_d) - log sig'(u0_d)] + sum_k sum_B s;  q(x) = 1/J.
With bf16=True the GEMM inputs (a, h1, h2, backward dY) and weights are
rounded to bfloat16 (round to nearest even) as the kernels round them;
accumulation is FP64 here (FP32 on the device).

Parity pinning (tests/test_flow_oracle.py): the inverse undoes the forward
(closed form), log J equals the log-determinant of the finite-difference
Jacobian of z -> x, and backprop() equals finite differences of the loss.
"""
from __future__ import annotations

import numpy as np

NH = 128      # hidden width
NI = 16       # padded input width (column 15 carries the layer-1 bias)
NO = 32       # padded output rows: s at 0..15, t at 16..31
SMAX = 4.0


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def masks(dim, nlayers):
    """Conditioning sets from the binary digits of the dimension index:
    layer k < nbits conditions on the dims whose bit k is 0, later layers on
    those whose bit (k - nbits) mod nbits is 1.  Any two dims differ in a bit,
    so some layer transforms one conditioned on the other (nlayers >= nbits)."""
    nbits = max(1, int(np.ceil(np.log2(dim))))
    out = []
    for k in range(nlayers):
        if k < nbits:
            a = np.array([((d >> k) & 1) == 0 for d in range(dim)])
        else:
            b = (k - nbits) % nbits
            a = np.array([((d >> b) & 1) == 1 for d in range(dim)])
        out.append(a)
    return out


def init_params(dim, nlayers=6, seed=0):
    """W1, W2 He-uniform; W3, b3 zero (identity map); b2 zero."""
    rng = np.random.default_rng(seed)
    P = []
    for k in range(nlayers):
        W1 = np.zeros((NH, NI))
        lim1 = np.sqrt(6.0 / (dim + 1))
        W1[:, :dim] = rng.uniform(-lim1, lim1, (NH, dim))
        W2 = rng.uniform(-np.sqrt(6.0 / NH), np.sqrt(6.0 / NH), (NH, NH))
        P.append(dict(W1=W1, W2=W2, b2=np.zeros(NH), W3=np.zeros((NO, NH)), b3=np.zeros(NO)))
    return P


def _logsigp(u):
    """log(sigmoid(u) (1 - sigmoid(u))), stable."""
    au = np.abs(u)
    return -au - 2.0 * np.log1p(np.exp(-au))


def _mlp(p, a, dim, rnd):
    W1 = rnd(p['W1'])
    W2 = rnd(p['W2'])
    W3 = rnd(p['W3'])
    h1p = rnd(a) @ W1.T
    h1 = np.maximum(h1p, 0.0)
    h2p = rnd(h1) @ W2.T + p['b2']
    h2 = np.maximum(h2p, 0.0)
    o = rnd(h2) @ W3.T + p['b3']
    return h1, h2, o


def _cond_input(u, A):
    n, dim = u.shape
    a = np.zeros((n, NI))
    a[:, :dim] = np.where(A[None, :], u, 0.0)
    a[:, NI - 1] = 1.0
    return a


def _st(o, A, dim):
    B = ~A
    s = SMAX * np.tanh(o[:, :dim] / SMAX) * B[None, :]
    t = o[:, 16:16 + dim] * B[None, :]
    return s, t


def forward(P, M, z, bf=True):
    """z -> (x, log J)."""
    rnd = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    z = np.asarray(z, np.float64)
    dim = z.shape[1]
    u = np.log(z) - np.log1p(-z)
    lj = -_logsigp(u).sum(1)
    for p, A in zip(P, M):
        _, _, o = _mlp(p, _cond_input(u, A), dim, rnd)
        s, t = _st(o, A, dim)
        u = u * np.exp(s) + t
        lj = lj + s.sum(1)
    lj = lj + _logsigp(u).sum(1)
    return 1.0 / (1.0 + np.exp(-u)), lj


def inverse(P, M, x, bf=True, keep=False):
    """x -> (z, log q(x)); keep=True also returns each layer's input u."""
    rnd = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    x = np.asarray(x, np.float64)
    dim = x.shape[1]
    u = np.log(x) - np.log1p(-x)
    lq = -_logsigp(u).sum(1)
    saved = [None] * len(P)
    for k in range(len(P) - 1, -1, -1):
        p, A = P[k], M[k]
        saved[k] = u.copy()
        _, _, o = _mlp(p, _cond_input(u, A), dim, rnd)
        s, t = _st(o, A, dim)
        u = (u - t) * np.exp(-s)
        lq = lq - s.sum(1)
    lq = lq + _logsigp(u).sum(1)
    z = 1.0 / (1.0 + np.exp(-u))
    return (z, lq, u, saved) if keep else (z, lq)


def backprop(P, M, x, c, bf=True):
    """Gradients of L = -sum_i c_i log q(x_i) with respect to every
    parameter (same dict layout as P), by explicit reverse-mode through the
    inverse pass.  With bf=True the backward GEMM inputs are rounded to
    bfloat16 as the kernels round them."""
    rnd = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    x = np.asarray(x, np.float64)
    c = np.asarray(c, np.float64)
    dim = x.shape[1]
    z, lq, u0, saved = inverse(P, M, x, bf, keep=True)
    # dL/du0 from the term sum_d log sig'(u0_d) of log q: d/du = 1 - 2 sig(u)
    G = -c[:, None] * (1.0 - 2.0 / (1.0 + np.exp(-u0)))
    grads = [None] * len(P)
    for k in range(len(P)):
        p, A = P[k], M[k]
        B = ~A
        uin = saved[k]
        a = _cond_input(uin, A)
        h1, h2, o = _mlp(p, a, dim, rnd)
        raw = o[:, :dim]
        th = np.tanh(raw / SMAX)
        s = SMAX * th * B[None, :]
        t = o[:, 16:16 + dim] * B[None, :]
        es = np.exp(-s)
        uout = (uin - t) * es
        gB = G * B[None, :]
        dt = -gB * es
        ds = (-gB * uout + c[:, None]) * B[None, :]
        draw = ds * (1.0 - th * th)
        do = np.zeros((len(x), NO))
        do[:, :dim] = draw
        do[:, 16:16 + dim] = dt
        W1, W2, W3 = rnd(p['W1']), rnd(p['W2']), rnd(p['W3'])
        dor = rnd(do)
        gW3 = dor.T @ rnd(h2)
        gb3 = do.sum(0)
        dh2 = (dor @ W3) * (h2 > 0)
        dh2r = rnd(dh2)
        gW2 = dh2r.T @ rnd(h1)
        gb2 = dh2.sum(0)
        dh1 = (dh2r @ W2) * (h1 > 0)
        dh1r = rnd(dh1)
        gW1 = dh1r.T @ rnd(a)
        da = dh1r @ W1
        Gin = G * es * B[None, :] + G * A[None, :]
        Gin = Gin + da[:, :dim] * A[None, :]
        G = Gin
        grads[k] = dict(W1=gW1, W2=gW2, b2=gb2, W3=gW3, b3=gb3)
    return grads, lq


def loss(P, M, x, c, bf=False):
    _, lq = inverse(P, M, x, bf)
    return -(np.asarray(c) * lq).sum()
