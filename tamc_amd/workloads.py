"""Synthetic instance generators for the benchmark and test workloads.

Every generator returns instance-file lines as three arrays ``(i, j, K)`` -- exactly what the
``i j K`` instance format (reference src/util.rs:59-109) would hold -- plus optional biases on the
diagonal (``i == j``).  Couplings are rounded to float32 before they are returned so that a
round trip through the text format is lossless.
"""
import numpy as np


def lattice_2d(L, rng=None, pm_j=True, ferro=False):
    """Periodic L x L square lattice; bonds (i -> right) and (i -> down) (row-major sites)."""
    n = L * L
    y, x = np.divmod(np.arange(n), L)
    i0 = np.repeat(np.arange(n), 2).astype(np.uint32)  # site s: (s, right of s), then (s, below s)
    j0 = np.stack([y * L + (x + 1) % L, ((y + 1) % L) * L + x], axis=1).ravel().astype(np.uint32)
    if ferro:
        K = -np.ones(len(i0))
    elif pm_j:
        K = rng.choice([-1.0, 1.0], size=len(i0))
    else:
        K = rng.standard_normal(len(i0)).astype(np.float32).astype(np.float64)
    return i0, j0, K, n


def lattice_3d(L, rng=None, pm_j=False, ferro=False):
    """Periodic L^3 cubic lattice with bonds to +x, +y, +z neighbours."""
    n = L ** 3
    idx = np.arange(n).reshape(L, L, L)
    i0, j0 = [], []
    for ax in range(3):
        nb = np.roll(idx, -1, axis=ax)
        i0.append(idx.ravel())
        j0.append(nb.ravel())
    i0 = np.concatenate(i0).astype(np.uint32)
    j0 = np.concatenate(j0).astype(np.uint32)
    # order lines by (i, then axis) like a row-major writer would
    order = np.lexsort((np.repeat(np.arange(3), n), i0))
    i0, j0 = i0[order], j0[order]
    if ferro:
        K = -np.ones(len(i0))
    elif pm_j:
        K = rng.choice([-1.0, 1.0], size=len(i0))
    else:
        K = rng.standard_normal(len(i0)).astype(np.float32).astype(np.float64)
    return i0, j0, K, n


def random_regular3(n, rng, gaussian=True, bias_sigma=0.0):
    """Random 3-regular simple graph by the configuration model with rejection."""
    assert n % 2 == 0 and n >= 4
    while True:
        stubs = np.repeat(np.arange(n), 3)
        rng.shuffle(stubs)
        a, b = stubs[0::2], stubs[1::2]
        if np.any(a == b):
            continue
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        key = lo.astype(np.int64) * n + hi
        if len(np.unique(key)) != len(key):
            continue
        break
    order = np.argsort(key)
    i0, j0 = lo[order].astype(np.uint32), hi[order].astype(np.uint32)
    if gaussian:
        K = rng.standard_normal(len(i0)).astype(np.float32).astype(np.float64)
    else:
        K = rng.choice([-1.0, 1.0], size=len(i0))
    if bias_sigma > 0:
        h = (rng.standard_normal(n) * bias_sigma).astype(np.float32).astype(np.float64)
        i0 = np.concatenate([i0, np.arange(n, dtype=np.uint32)])
        j0 = np.concatenate([j0, np.arange(n, dtype=np.uint32)])
        K = np.concatenate([K, h])
    return i0, j0, K, n


def qubo_sparse(n, density, rng, qmax=100, diagonal=True):
    """Upper-triangular QUBO: each pair present with probability `density`, Q ~ U{-qmax..qmax}."""
    iu, ju = np.triu_indices(n, k=1)
    keep = rng.random(len(iu)) < density
    i0, j0 = iu[keep].astype(np.uint32), ju[keep].astype(np.uint32)
    K = rng.integers(-qmax, qmax + 1, size=len(i0)).astype(np.float64)
    if diagonal:
        i0 = np.concatenate([i0, np.arange(n, dtype=np.uint32)])
        j0 = np.concatenate([j0, np.arange(n, dtype=np.uint32)])
        K = np.concatenate([K, rng.integers(-qmax, qmax + 1, size=n).astype(np.float64)])
    return i0, j0, K, n


def write_instance_file(path, i, j, K):
    """Write lines in fixed-point decimal (never exponent notation: the reference parser stops at 'e')."""
    with open(path, "w") as f:
        for a, b, k in zip(i, j, K):
            f.write("%d %d %s\n" % (a, b, np.format_float_positional(np.float32(k), unique=True, trim="0")))


def config3_instances(count=256, seed=103, nmin=200, nmax=4000):
    """Batched PT-ICM workload: random 3-regular graphs, N ~ U{nmin..nmax} even, J ~ N(0,1), h ~ N(0, 0.1)."""
    out = []
    for k in range(count):
        rng = np.random.default_rng(seed + k)
        n = int(rng.integers(nmin // 2, nmax // 2 + 1)) * 2
        out.append(random_regular3(n, rng, gaussian=True, bias_sigma=0.1))
    return out


def serving_requests(count=2000, rate=200.0, seed=105):
    """Serving workload: Poisson arrivals at `rate` per second; each request is, with equal
    probability, a periodic 2D lattice L in [8,128], a periodic 3D lattice L in [4,24] or a random
    3-regular graph N in [64,16384] (even), with +-J or N(0,1) couplings (probability 1/2 each).
    Returns a list of (arrival_s, kind, size, (i, j, K, n))."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        kind = int(rng.integers(0, 3))
        pm = bool(rng.integers(0, 2))
        sub = np.random.default_rng(int(rng.integers(0, 2**63 - 1)))
        if kind == 0:
            L = int(rng.integers(8, 129))
            lines = lattice_2d(L, sub, pm_j=pm)
            out.append((t, "2d", L, lines))
        elif kind == 1:
            L = int(rng.integers(4, 25))
            lines = lattice_3d(L, sub, pm_j=pm)
            out.append((t, "3d", L, lines))
        else:
            n = int(rng.integers(32, 8193)) * 2
            lines = random_regular3(n, sub, gaussian=not pm)
            out.append((t, "rr3", n, lines))
    return out
