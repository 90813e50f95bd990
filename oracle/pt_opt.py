"""Restatement of one temperature-set update of pt_optimize_beta (src/bin/tamc-pt-opt.rs:57-167)
and of the tamc-core helpers it uses (tamc-core/src/util.rs:6-138), in float32 with the
reference's operation order.  logf/expf/log10f come from the C library, as f32::ln/exp/log10 do.

Test infrastructure: the product package never imports this module.

Parity unpinned: ndarray's `sum` (restated as its eight-way unrolled fold) and the `interp` crate's
linear interpolation (restated as plain piecewise-linear interpolation, clamped at the ends) are
third-party code absent here, and no reference test or vector covers them.
"""
import ctypes
import ctypes.util

import numpy as np

_m = ctypes.CDLL(ctypes.util.find_library("m"))
for _name in ("logf", "expf", "log10f", "sqrtf"):
    getattr(_m, _name).restype = ctypes.c_float
    getattr(_m, _name).argtypes = [ctypes.c_float]

f32 = np.float32


def logf(x):
    return f32(_m.logf(float(x)))


def expf(x):
    return f32(_m.expf(float(x)))


def log10f(x):
    return f32(_m.log10f(float(x)))


def unrolled_sum(xs):
    """ndarray's sum of a contiguous f32 array: eight partial sums, combined pairwise, then the tail."""
    xs = [f32(x) for x in xs]
    p = [f32(0.0)] * 8
    k = 0
    while len(xs) - k >= 8:
        for a in range(8):
            p[a] = f32(p[a] + xs[k + a])
        k += 8
    acc = f32(0.0)
    for a in range(4):
        acc = f32(acc + f32(p[a] + p[a + 4]))
    for x in xs[k:]:
        acc = f32(acc + x)
    return acc


def stepwise_weights(x):
    """StepwiseMeasure::new (tamc-core/src/util.rs:24-43)."""
    n = len(x)
    dx = [f32(x[i + 1] - x[i]) for i in range(n - 1)]
    w = [f32(0.0)] * n
    w[0] = f32(dx[0] / f32(2.0))
    for i in range(1, n - 1):
        w[i] = f32(f32(dx[i - 1] + dx[i]) / f32(2.0))
    w[n - 1] = f32(dx[n - 2] / f32(2.0))
    return w


def finite_differences(x, y):
    """finite_differences (tamc-core/src/util.rs:6-22)."""
    n = len(x)
    dx = [f32(x[i + 1] - x[i]) for i in range(n - 1)]
    dy = [f32(y[i + 1] - y[i]) for i in range(n - 1)]
    d = [f32(0.0)] * n
    d[0] = f32(dy[0] / dx[0])
    d[n - 1] = f32(dy[-1] / dx[-1])
    for i in range(1, n - 1):
        a = f32(f32(dy[i - 1] * dx[i]) / dx[i - 1])
        b = f32(f32(dy[i] * dx[i - 1]) / dx[i])
        d[i] = f32(f32(a + b) / f32(dx[i - 1] + dx[i]))
    return d


def interp(x, y, xp):
    xp = f32(xp)
    if xp <= x[0]:
        return f32(y[0])
    if xp >= x[-1]:
        return f32(y[-1])
    i = max(k for k in range(len(x) - 1) if x[k] <= xp)
    return f32(y[i] + f32(f32(f32(xp - x[i]) * f32(y[i + 1] - y[i])) / f32(x[i + 1] - x[i])))


def monotonic_bisection(f, y0, xmin, xmax, tol=f32(1e-6), iters_max=1000000):
    """monotonic_bisection (tamc-core/src/util.rs:62-92): ("root" | "lo" | "up" | "failed", x)."""
    assert xmin < xmax
    x0, x1 = f32(xmin), f32(xmax)
    for _ in range(iters_max):
        xmid = f32(f32(x0 + x1) / f32(2.0))
        y = f(xmid)
        if abs(f32(y - y0)) < tol:
            return "root", xmid
        if abs(f32(xmid - xmin)) < tol:
            return "lo", xmid
        if abs(f32(xmid - xmax)) < tol:
            return "up", xmid
        if y < y0:
            x0 = xmid
        else:
            x1 = xmid
    return "failed", None


def bisect_array(f, roots, lo_idx, ytgt, xmin, xmax):
    """bisect_array (tamc-core/src/util.rs:96-125) on roots[lo_idx : lo_idx + len(ytgt)]."""
    n = len(ytgt)
    if n == 0:
        return
    if n == 1:
        kind, x = monotonic_bisection(f, ytgt[0], xmin, xmax)
        if kind != "root":
            raise ValueError("bisection for %s in [%s, %s] ended: %s" % (ytgt[0], xmin, xmax, kind))
        roots[lo_idx] = x
        return
    mid = n // 2
    bisect_array(f, roots, lo_idx + mid, ytgt[mid:mid + 1], xmin, xmax)
    xmid = roots[lo_idx + mid]
    bisect_array(f, roots, lo_idx, ytgt[:mid], xmin, xmid)
    bisect_array(f, roots, lo_idx + mid + 1, ytgt[mid + 1:], xmid, xmax)


def monotonic_divisions(f, m):
    """monotonic_divisions (tamc-core/src/util.rs:127-138)."""
    parts = [f32(i / m) for i in range(1, m)]
    x = [f32(0.0)] * (m + 1)
    x[m] = f32(1.0)
    bisect_array(f, x, 1, parts, f32(0.0), f32(1.0))
    return x


def update(beta_vec, momentum, dif_hists, round_trips, num_sweeps, num_chains, alpha, m):
    """One iteration's update.  dif_hists[k] = diffusion histograms [nt][2] summed over chains,
    round_trips[k] = round trips summed over chains.  Returns (new_betas, new_momentum, err, taus)
    or None when the eta normalisation is below f32 epsilon (the reference skips the update)."""
    beta_vec = [f32(b) for b in beta_vec]
    nt = len(beta_vec)
    alpha, m = f32(alpha), f32(m)
    weights = stepwise_weights(beta_vec)
    difs = [f32(beta_vec[i + 1] - beta_vec[i]) for i in range(nt - 1)]
    taus, ddifs = [], []
    for h, rts in zip(dif_hists, round_trips):
        h = np.asarray(h, np.uint64)
        rt_per_rep = f32(f32(int(rts)) / f32(num_chains * nt))
        taus.append(f32(0.0) if int(rts) == 0 else f32(f32(num_sweeps) / rt_per_rep))
        probs = []
        for b in range(nt):
            a0, a1 = f32(int(h[b, 0]) + 1), f32(int(h[b, 1]) + 1)
            probs.append(f32(a1 / f32(a0 + a1)))
        ddifs.append(finite_differences(beta_vec, probs))
    sum_tau = unrolled_sum(taus)
    w = [f32(0.0)] * nt
    for tau, d in zip(taus, ddifs):
        w = [f32(w[i] + f32(tau * d[i])) for i in range(nt)]
    with np.errstate(invalid="ignore", divide="ignore"):
        w = [f32(x / sum_tau) for x in w]
    # f32::max returns the other operand for a NaN (0 / 0 when no instance made a round trip)
    eta = [f32(np.sqrt(np.fmax(f32(w[i] / weights[i]), f32(0.0)))) for i in range(nt)]
    eta_w = [f32(f32(f32(eta[i] + eta[i + 1]) / f32(2.0)) * difs[i]) for i in range(nt - 1)]
    z = unrolled_sum(eta_w)
    if z < np.finfo(np.float32).eps:
        return None
    cdf, acc = [], f32(0.0)
    for x in eta_w:
        cdf.append(acc)
        acc = f32(acc + f32(x / z))
    cdf.append(f32(1.0))
    bmin, bmax = beta_vec[0], beta_vec[-1]

    def eta_fn(x):
        return interp(beta_vec, cdf, f32(bmin + f32(x * f32(bmax - bmin))))

    xd = monotonic_divisions(eta_fn, nt - 1)
    calc = [f32(f32(x * f32(bmax - bmin)) + bmin) for x in xd]
    mom = [expf(f32(f32(m * logf(bp)) + f32(f32(f32(1.0) - m) * logf(b1)))) for bp, b1 in zip(momentum, calc)]
    out = list(calc)
    err = f32(0.0)
    for i in range(1, nt - 1):
        b = expf(f32(f32(alpha * logf(mom[i])) + f32(f32(f32(1.0) - alpha) * logf(beta_vec[i]))))
        err = f32(err + abs(f32(log10f(out[i]) - log10f(beta_vec[i]))))
        out[i] = b
    err = f32(err / f32(nt))
    return out, mom, err, taus
