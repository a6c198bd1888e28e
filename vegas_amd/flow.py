"""Learned map: an affine-coupling normalizing flow on [0,1]^D trained on
samples from vegas itself (libvgx ``vgx_flow_*``; BF16 MFMA kernels, FP32
accumulation and log-determinant).

The flow x = F(z), z uniform on [0,1]^D, replaces the separable vegas grid
for integrands whose peaks do not line up with the axes.  It is trained by
minimising the variance of f/q, the per-sample variance of the importance
sampling estimate: on samples x_i with sampling weights w_i (vegas' own
weights, so that sum_i w_i f(x_i) estimates the integral) the loss
sum_i w_i f_i^2 / q(x_i) estimates  int f^2/q dx = Var_q(f/q) + I^2.  The
vegas grid stays the fallback and the arbiter: :meth:`FlowMap.accept` keeps
the flow only when its measured per-sample variance is lower.

Layout of the FP32 parameter vector (per layer): W1[128][16] (column 15 is
the first bias), W2[128][128], b2[128], W3[32][128], b3[32]; rows 0..15 of
W3 give the log-scales s (through 4 tanh(./4)), rows 16..31 the shifts t.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _lib
from ._ravg import GVar

NH, NI, NO = 128, 16, 32
PLAYER = NH * NI + NH * NH + NH + NO * NH + NO
OFF_W1, OFF_W2 = 0, NH * NI
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NO * NH


def _bind(L):
    if getattr(L, '_vgx_flow_bound', False):
        return L
    vp, i64, f32 = C.c_void_p, C.c_int64, C.c_float
    L.vgx_flow_create.argtypes = [C.c_int, C.c_int, vp, C.POINTER(vp)]
    L.vgx_flow_destroy.argtypes = [vp]
    L.vgx_flow_nparams.argtypes = [vp, C.POINTER(i64)]
    L.vgx_flow_set_params.argtypes = [vp, vp, vp]
    L.vgx_flow_get_params.argtypes = [vp, vp, vp, vp]
    L.vgx_flow_sample.argtypes = [vp, C.POINTER(_lib.VgxIntegrand), i64, C.c_uint64, C.c_uint64, C.c_uint32,
                                  vp, vp, vp, vp, vp]
    L.vgx_flow_logq.argtypes = [vp, vp, i64, vp, vp, vp]
    L.vgx_flow_grad.argtypes = [vp, vp, vp, vp, i64, C.c_int, vp]
    L.vgx_flow_adam.argtypes = [vp, f32, f32, f32, f32, f32, vp]
    L.vgx_flow_eval.argtypes = [vp, C.POINTER(_lib.VgxIntegrand), vp, i64, vp, vp]
    L.vgx_flow_mfma_probe.argtypes = [vp, vp, vp, vp]
    L._vgx_flow_bound = True
    return L


def masks(dim, nlayers):
    """Conditioning sets from the binary digits of the dimension index:
    layer k < nbits conditions on the dims whose bit k is 0, later layers on
    those whose bit (k - nbits) mod nbits is 1, so every pair of dims is
    split by some layer."""
    nbits = max(1, int(math.ceil(math.log2(dim)))) if dim > 1 else 1
    out = []
    for k in range(nlayers):
        if k < nbits:
            a = [((d >> k) & 1) == 0 for d in range(dim)]
        else:
            b = (k - nbits) % nbits
            a = [((d >> b) & 1) == 1 for d in range(dim)]
        out.append(np.array(a))
    return out


def init_params(dim, nlayers=6, seed=0):
    """Flat FP32 parameters: W1, W2 He-uniform, W3 and b3 zero (the
    identity map), biases zero."""
    rng = np.random.default_rng(seed)
    P = np.zeros((nlayers, PLAYER), np.float32)
    for k in range(nlayers):
        W1 = np.zeros((NH, NI))
        lim1 = math.sqrt(6.0 / (dim + 1))
        W1[:, :dim] = rng.uniform(-lim1, lim1, (NH, dim))
        W2 = rng.uniform(-math.sqrt(6.0 / NH), math.sqrt(6.0 / NH), (NH, NH))
        P[k, OFF_W1:OFF_W2] = W1.reshape(-1)
        P[k, OFF_W2:OFF_B2] = W2.reshape(-1)
    return P


def unpack(P):
    """Per-layer dicts (W1, W2, b2, W3, b3) of a flat parameter array."""
    P = np.asarray(P, np.float64).reshape(-1, PLAYER)
    return [dict(W1=p[OFF_W1:OFF_W2].reshape(NH, NI), W2=p[OFF_W2:OFF_B2].reshape(NH, NH),
                 b2=p[OFF_B2:OFF_W3].copy(), W3=p[OFF_W3:OFF_B3].reshape(NO, NH), b3=p[OFF_B3:].copy())
            for p in P]


def pack(layers):
    out = np.zeros((len(layers), PLAYER), np.float32)
    for k, p in enumerate(layers):
        out[k, OFF_W1:OFF_W2] = np.asarray(p['W1']).reshape(-1)
        out[k, OFF_W2:OFF_B2] = np.asarray(p['W2']).reshape(-1)
        out[k, OFF_B2:OFF_W3] = p['b2']
        out[k, OFF_W3:OFF_B3] = np.asarray(p['W3']).reshape(-1)
        out[k, OFF_B3:] = p['b3']
    return out


class CouplingFlow:
    """The flow's parameters on the device and its kernels."""

    def __init__(self, dim, nlayers=6, seed=0, params=None):
        self._torch = _lib.torch_cuda()
        if not 1 <= dim <= 15:
            raise ValueError('flow dimension must be 1..15')
        self.dim, self.nlayers = int(dim), int(nlayers)
        self.masks = masks(self.dim, self.nlayers)
        am = (C.c_uint32 * self.nlayers)(*[int(sum(1 << d for d in range(self.dim) if m[d]))
                                           for m in self.masks])
        self._L = _bind(_lib.lib())
        h = C.c_void_p()
        _lib.check(self._L.vgx_flow_create(self.dim, self.nlayers, C.cast(am, C.c_void_p), C.byref(h)),
                   'vgx_flow_create')
        self._h = h
        self.set_params(init_params(self.dim, self.nlayers, seed) if params is None else params)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                self._L.vgx_flow_destroy(h)
            except Exception:
                pass

    # ---------------------------------------------------------- parameters
    def set_params(self, P):
        P = np.ascontiguousarray(np.asarray(P, np.float32).reshape(self.nlayers, PLAYER))
        _lib.check(self._L.vgx_flow_set_params(self._h, P.ctypes.data, _lib.stream_ptr()), 'vgx_flow_set_params')
        self.steps = 0

    def params(self):
        P = np.zeros((self.nlayers, PLAYER), np.float32)
        _lib.check(self._L.vgx_flow_get_params(self._h, P.ctypes.data, None, _lib.stream_ptr()),
                   'vgx_flow_get_params')
        return P

    def last_grad(self):
        G = np.zeros((self.nlayers, PLAYER), np.float32)
        _lib.check(self._L.vgx_flow_get_params(self._h, None, G.ctypes.data, _lib.stream_ptr()),
                   'vgx_flow_get_params')
        return G

    # ------------------------------------------------------------- kernels
    def padded(self, x):
        """Points as the device float [n][16] the kernels read (returned as
        is when already in that form)."""
        torch = self._torch
        x = torch.as_tensor(x, device='cuda')
        if x.dtype == torch.float32 and x.dim() == 2 and x.shape[1] == 16 and x.is_contiguous():
            return x
        out = torch.zeros((x.shape[0], 16), dtype=torch.float32, device='cuda')
        out[:, :self.dim] = x.to(torch.float32)
        return out

    def logq(self, x):
        """log q(x) and logit(z) of the preimages, for device/host points x."""
        torch = self._torch
        xp = self.padded(x)
        n = xp.shape[0]
        lq = torch.empty(n, dtype=torch.float32, device='cuda')
        u0 = torch.empty((n, 16), dtype=torch.float32, device='cuda')
        _lib.check(self._L.vgx_flow_logq(self._h, _lib.ptr(xp), n, _lib.ptr(lq), _lib.ptr(u0), _lib.stream_ptr()),
                   'vgx_flow_logq')
        return lq, u0[:, :self.dim]

    def grad(self, x, f, logw, loss='variance'):
        """Device gradient of the loss ('variance', 'kl', or 'squared': the
        variance loss for the flow's own samples, given f J as f and log w =
        0) on points x (float [n][16]), values f and log sampling weights
        logw."""
        torch = self._torch
        n = x.shape[0]
        f = torch.as_tensor(f, dtype=torch.float32, device='cuda').contiguous()
        w = torch.as_tensor(logw, dtype=torch.float32, device='cuda').contiguous()
        _lib.check(self._L.vgx_flow_grad(self._h, _lib.ptr(x), _lib.ptr(f), _lib.ptr(w), n,
                                         {'variance': 0, 'kl': 1, 'squared': 2}[loss], _lib.stream_ptr()),
                   'vgx_flow_grad')

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, clip=0.0):
        _lib.check(self._L.vgx_flow_adam(self._h, lr, beta1, beta2, eps, clip, _lib.stream_ptr()), 'vgx_flow_adam')
        self.steps += 1

    def eval_fcn(self, fcn, x):
        """f(x) of a device integrand at device points x float [n][16]."""
        torch = self._torch
        vf, keep = fcn.vgx_integrand(torch, torch.device('cuda'))
        n = x.shape[0]
        out = torch.empty(n, dtype=torch.float32, device='cuda')
        _lib.check(self._L.vgx_flow_eval(self._h, C.byref(vf), _lib.ptr(x), n, _lib.ptr(out), _lib.stream_ptr()),
                   'vgx_flow_eval')
        return out

    def sample(self, fcn, n, seed=1, iteration=0, base=0, keep_points=False, status=None):
        """sum f J and sum (f J)^2 over n flow samples; with keep_points also
        the points (device float [n][16]) and f J (device float [n]).  Given a
        device int32 ``status`` (with keep_points), nothing is read back: the
        sums are None, the call does not wait for the GPU, and a NaN integrand
        shows as VGX_ENAN in ``status`` for the caller to check."""
        torch = self._torch
        vf, keep = fcn.vgx_integrand(torch, torch.device('cuda'))
        lazy = status is not None
        if lazy and not keep_points:
            raise ValueError('a lazy sample needs keep_points')
        sums = None if lazy else (C.c_double * 2)()
        st = status if lazy else torch.zeros(1, dtype=torch.int32, device='cuda')
        xo = fo = None
        if keep_points:
            xo = torch.empty((int(n), 16), dtype=torch.float32, device='cuda')
            fo = torch.empty(int(n), dtype=torch.float32, device='cuda')
        _lib.check(self._L.vgx_flow_sample(self._h, C.byref(vf), int(n), int(base), int(seed), int(iteration),
                                           None if lazy else C.cast(sums, C.c_void_p), _lib.ptr(xo),
                                           _lib.ptr(fo), _lib.ptr(st), _lib.stream_ptr()), 'vgx_flow_sample')
        if lazy:
            return None, None, xo, fo
        if int(st.item()) == _lib.VGX_ENAN:
            raise ValueError('integrand evaluates to nan')
        return (sums[0], sums[1], xo, fo) if keep_points else (sums[0], sums[1])


def vegas_training_data(integ, fcn, flow, n):
    """n samples (x, f, log w) from the current vegas map of ``integ``:
    whole iterations of its stratified sampler, weights divided by the
    number of iterations (so sum w f estimates the integral)."""
    torch = _lib.torch_cuda()
    xs, ws = [], []
    got, nit = 0, 0
    while got < n:
        for x, w in integ._device_batches():
            xs.append(x)
            ws.append(w)
            got += x.shape[0]
        nit += 1
    x = torch.cat(xs)[:n]
    w = torch.cat(ws)[:n] / nit
    xp = flow.padded(x)
    f = flow.eval_fcn(fcn, xp)
    return xp, f, torch.log(w).to(torch.float32)


class FlowMap:
    """Trains a :class:`CouplingFlow` for an integrand and integrates with it.

    train(): Adam on the variance loss over batches of vegas samples, then
    (optionally) over samples of the flow itself.  integrate(): the plain
    importance-sampling estimate from flow samples until the relative error
    reaches rtol.  accept(): the flow's per-sample variance against the
    vegas grid's (the fallback)."""

    def __init__(self, dim, nlayers=6, seed=0):
        self.flow = CouplingFlow(dim, nlayers, seed)
        self.seed = seed
        self.history = []

    def train(self, fcn, x, f, logw, epochs=4, batch=1000000, lr=1e-3, clip=0.0, self_batches=0,
              self_batch=1000000, loss='variance'):
        torch = _lib.torch_cuda()
        n = x.shape[0] if x is not None else 0
        t0 = time.perf_counter()
        # batch order from a generator seeded by the flow's seed: the
        # training, and so the learned map, repeat for a given seed
        gen = torch.Generator(device='cuda')
        gen.manual_seed(int(self.seed) * 1000003 + 17)
        for ep in range(epochs if n else 0):
            perm = torch.randperm(n, device='cuda', generator=gen)
            for b0 in range(0, n, batch):
                idx = perm[b0:b0 + batch]
                self.flow.grad(x[idx].contiguous(), f[idx], logw[idx], loss)
                self.flow.adam(lr=lr, clip=clip)
        status = torch.zeros(1, dtype=torch.int32, device='cuda')
        zero_w = torch.zeros(self_batch, dtype=torch.float32, device='cuda') if self_batches else None
        for it in range(self_batches):
            # the flow's own samples (weights 1/(n q) = J/n): the loss
            # weights w f^2 / q and w |f| are (f J)^2 / n and |f J| / n, so
            # the sampler's f J is passed with log w = 0 and no log q or
            # second integrand evaluation is needed; no read-back inside the
            # loop (NaN is checked once at the end)
            self._self_it = getattr(self, '_self_it', 0) + 1
            _, _, xo, fo = self.flow.sample(fcn, self_batch, seed=self.seed + 7919, iteration=1000 + self._self_it,
                                            keep_points=True, status=status)
            self.flow.grad(xo, fo, zero_w, 'squared' if loss == 'variance' else 'kl')
            self.flow.adam(lr=lr, clip=clip)
        torch.cuda.synchronize()
        if self_batches and int(status.item()) == _lib.VGX_ENAN:
            raise ValueError('integrand evaluates to nan')
        self.train_time = time.perf_counter() - t0
        return self

    def fit(self, fcn, integ, ntrain=20000000, batch=250000, kl_epochs=3, self_kl=200, self_var=200,
            lr=1e-3, clip=1.0, rtol=None, chunk=100, max_var=3000, probe=4000000):
        """The training recipe: ntrain samples from the adapted vegas map of
        ``integ``; Adam on the forward KL loss over them (stable when the
        samples see only part of the peak), then on the flow's own samples
        (KL, then the variance loss at a lower rate).  Batches of 2.5e5: on
        the 9-D ridge more, smaller Adam steps reach a lower variance per
        unit of training time than batches of 1e6.  The step counts are the
        ones that minimised training plus sampling time to 1e-4 there (3
        epochs, 200 + 200 self-training steps: 1.7 s, against 2.2 s for 6
        epochs and 300 + 150).

        With ``rtol`` the variance-loss phase continues past self_var steps,
        ``chunk`` steps at a time (at most max_var in all): after each, the
        per-sample variance on ``probe`` samples gives the sampling time
        still needed for rtol, and training stops after two chunks in a row
        that did not lower the projected time to the target (training so far
        plus that sampling time), and the weights of the best projection
        are restored (the variance loss is noisy on heavy-tailed weights: a
        chunk can undo the previous ones).  On the narrower 8-D ridge the
        fixed 200 steps leave 2-3 times the variance that 1000 more remove.
        The flow's weights then depend on these timings (not on the seed
        alone).""" 
        x, fx, lw = vegas_training_data(integ, fcn, self.flow, int(ntrain))
        self.train(fcn, x, fx, lw, epochs=kl_epochs, batch=batch, lr=lr, clip=clip, loss='kl')
        del x, fx, lw
        if self_kl:
            self.train(fcn, None, None, None, epochs=0, self_batches=self_kl, self_batch=batch, lr=lr, clip=clip,
                       loss='kl')
        if self_var:
            self.train(fcn, None, None, None, epochs=0, self_batches=self_var, self_batch=batch, lr=lr * 0.3,
                       clip=clip, loss='variance')
        if rtol is None:
            return self

        def remaining():
            t0 = time.perf_counter()
            mean, var = self.per_sample_variance(fcn, n=probe, iteration=9000 + len(self.history))
            per_sample = (time.perf_counter() - t0) / probe
            need = var / max((rtol * mean) ** 2, 1e-300)
            return need * per_sample
        torch = _lib.torch_cuda()
        t_start = time.perf_counter()
        best = remaining()  # projected time from here: further training plus sampling
        best_p = self.flow.params()
        done, stale = self_var, 0
        while done < max_var and stale < 2:
            self.train(fcn, None, None, None, epochs=0, self_batches=chunk, self_batch=batch, lr=lr * 0.3,
                       clip=clip, loss='variance')
            now = remaining()
            torch.cuda.synchronize()
            total = (time.perf_counter() - t_start) + now
            done += chunk
            self.history.append((done, now))
            if total < best:
                best, stale = total, 0
                best_p = self.flow.params()
            else:
                stale += 1
        if stale:
            # the last chunks did not improve on the best projection: its
            # weights (Adam's state restarts with them)
            self.flow.set_params(best_p)
        return self

    def per_sample_variance(self, fcn, n=1000000, iteration=0):
        s1, s2 = self.flow.sample(fcn, n, seed=self.seed + 1, iteration=iteration)
        mean = s1 / n
        return mean, max(s2 / n - mean * mean, 0.0)

    def accept(self, fcn, vegas_var_per_sample, n=1000000):
        """Use the flow only when its per-sample variance is lower than the
        vegas grid's (var of the vegas estimate times its sample count)."""
        mean, var = self.per_sample_variance(fcn, n)
        self.accepted = var < vegas_var_per_sample
        self.flow_var_per_sample = var
        return self.accepted

    def integrate(self, fcn, rtol=1e-4, batch=10000000, max_samples=10 ** 11, seed=None, iteration=0):
        """Flow importance sampling until sdev < rtol |mean|; returns a GVar
        and the number of samples."""
        seed = self.seed + 2 if seed is None else seed
        s1 = s2 = 0.0
        n = 0
        base = 0
        while n < max_samples:
            a, b = self.flow.sample(fcn, batch, seed=seed, iteration=iteration, base=base)
            s1 += a
            s2 += b
            n += batch
            base += batch
            mean = s1 / n
            var = max(s2 / n - mean * mean, 0.0) / n
            if n >= 2 * batch and math.sqrt(var) < rtol * abs(mean):
                break
        return GVar(mean, math.sqrt(var)), n
