"""Exact thermodynamics of the nearest-neighbour Ising ferromagnet on a periodic m x n lattice
(H = -sum_<ij> s_i s_j), from Kaufman's partition function (B. Kaufman, Phys. Rev. 76, 1232
(1949); the form used by P. D. Beale, Phys. Rev. Lett. 76, 78 (1996)):

    Z = 1/2 (2 sinh 2K)^(mn/2) (Z1 + Z2 + Z3 + Z4),   K = beta,
    Z1 = prod_{r<n} 2 cosh(m g_{2r+1} / 2),   Z2 = prod_{r<n} 2 sinh(m g_{2r+1} / 2),
    Z3 = prod_{r<n} 2 cosh(m g_{2r} / 2),     Z4 = prod_{r<n} 2 sinh(m g_{2r} / 2),
    cosh g_k = cosh 2K coth 2K - cos(pi k / n)  (g_k > 0, k > 0),  g_0 = 2K + ln tanh K.

<E>(beta) = -d ln Z / d beta is taken by a central difference of ln Z, which is summed in logs
with signs (Z4 is negative below the critical coupling when g_0 < 0).  Test helper only.
"""
import math

import numpy as np


def _log_prod(vals):
    """log|prod| and sign of a product of reals."""
    sgn, lg = 1.0, 0.0
    for v in vals:
        if v == 0.0:
            return 0.0, -math.inf
        sgn *= math.copysign(1.0, v)
        lg += math.log(abs(v))
    return sgn, lg


def log_z(K, m, n):
    g = [2 * K + math.log(math.tanh(K))]
    c = math.cosh(2 * K) / math.tanh(2 * K)
    for k in range(1, 2 * n):
        g.append(math.acosh(c - math.cos(math.pi * k / n)))
    odd = [g[2 * r + 1] for r in range(n)]
    even = [g[2 * r] for r in range(n)]
    terms = [_log_prod([2 * math.cosh(m * x / 2) for x in odd]), _log_prod([2 * math.sinh(m * x / 2) for x in odd]),
             _log_prod([2 * math.cosh(m * x / 2) for x in even]), _log_prod([2 * math.sinh(m * x / 2) for x in even])]
    top = max(lg for _, lg in terms)
    s = sum(sg * math.exp(lg - top) for sg, lg in terms)
    return math.log(0.5) + (m * n / 2) * math.log(2 * math.sinh(2 * K)) + top + math.log(s)


def mean_energy(beta, m, n, h=1e-5):
    return -(log_z(beta + h, m, n) - log_z(beta - h, m, n)) / (2 * h)


def brute_force_mean_energy(beta, m, n):
    """Enumeration over all 2^(mn) states (small lattices only)."""
    N = m * n
    idx = np.arange(N).reshape(m, n)
    right = np.roll(idx, -1, axis=1).ravel()
    down = np.roll(idx, -1, axis=0).ravel()
    states = 1 - 2 * ((np.arange(2 ** N)[:, None] >> np.arange(N)) & 1).astype(np.int8)
    E = -(states * states[:, right]).sum(axis=1) - (states * states[:, down]).sum(axis=1)
    E = E.astype(np.float64)
    w = np.exp(-beta * (E - E.min()))
    return float((w * E).sum() / w.sum())
