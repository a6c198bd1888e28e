"""FP64 restatement of the multi-component per-hcube sums (test
infrastructure only; only tests may import it).

src/vegas/_vegas.pyx:2136-2183 with fcn_size = m > 1: for every hcube, in
point order, sum_wf[s] = sum w f_s; then per point the deviations
dwf[s] = w f_s - sum_wf[s]/n with the EPSILON guard (a guarded deviation
adds (EPSILON |mean|)^2 to its diagonal entry and is zeroed before it enters
sum_dwf or any off-diagonal product); per hcube mean[s] += sum_wf[s] +
sum_dwf[s] and var[s,t] += (n sum_dwf2[s,t] - sum_dwf[s] sum_dwf[t])/(n-1)
(t <= s with correlate_integrals, else the diagonal only, stored in column
0 as the reference stores it).  Pure-Python loops: small inputs only.

Parity pinning (tests/test_multi_oracle.py): away from the guard the per
hcube terms are n times numpy's sample covariance of w f and the mean is
sum w f (textbook identities); a component that is a multiple of another
gives the scaled variance and covariance; m = 1 matches the scalar rule of
the C oracle on the same points; a constant component takes the guard.
"""
from __future__ import annotations

import numpy as np

EPSILON = np.finfo(float).eps * 1e4  # pyx:36


def iteration_sums(hcube, wgt, fx, correlate=True):
    """Mean (m,) and covariance (m, m) (correlate) or variances (m,) of one
    iteration from points sorted by hcube (pyx:2136-2183, 2199-2204)."""
    hcube = np.asarray(hcube)
    wgt = np.asarray(wgt, np.float64)
    fx = np.asarray(fx, np.float64)
    N, m = fx.shape
    mean = np.zeros(m)
    var = np.zeros((m, m))
    j = 0
    while j < N:
        i = hcube[j]
        sum_wf = np.zeros(m)
        sum_dwf = np.zeros(m)
        sum_dwf2 = np.zeros((m, m))
        neval = 0
        jtmp = j
        while jtmp < N and hcube[jtmp] == i:
            for s in range(m):
                sum_wf[s] += wgt[jtmp] * fx[jtmp, s]
            jtmp += 1
            neval += 1
        while j < N and hcube[j] == i:
            dwf = np.zeros(m)
            for s in range(m):
                dwf[s] = wgt[j] * fx[j, s] - sum_wf[s] / neval
                g = EPSILON * abs(sum_wf[s] / neval)
                if abs(dwf[s]) < g:
                    dwf[s] = g
                    sum_dwf2[s, s if correlate else 0] += dwf[s] ** 2
                    dwf[s] = 0.
                else:
                    sum_dwf2[s, s if correlate else 0] += dwf[s] ** 2
                sum_dwf[s] += dwf[s]
                if correlate:
                    for t in range(s):
                        sum_dwf2[s, t] += dwf[s] * dwf[t]
            j += 1
        for s in range(m):
            mean[s] += sum_wf[s] + sum_dwf[s]
            if correlate:
                for t in range(s + 1):
                    var[s, t] += (neval * sum_dwf2[s, t] - sum_dwf[s] * sum_dwf[t]) / (neval - 1.)
            else:
                var[s, 0] += (neval * sum_dwf2[s, 0] - sum_dwf[s] * sum_dwf[s]) / (neval - 1.)
    if correlate:
        for s in range(m):
            for t in range(s):
                var[t, s] = var[s, t]
        return mean, var
    return mean, var[:, 0].copy()
