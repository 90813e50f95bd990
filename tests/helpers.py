"""Shared builders: the same instance lines go to the oracle (CPU restatement) and the product."""
import numpy as np

import oracle as O
from tamc_amd import workloads as W


def make_pair(kind, seed=0, qubo=False, **kw):
    """Return (oracle instance, product instance, lines) for a named small workload."""
    from tamc_amd.api import BqmIsingInstance
    rng = np.random.default_rng(seed)
    if kind == "2d_pm":
        i, j, K, n = W.lattice_2d(kw.get("L", 8), rng, pm_j=True)
    elif kind == "2d_ferro":
        i, j, K, n = W.lattice_2d(kw.get("L", 8), ferro=True)
    elif kind == "2d_gauss":
        i, j, K, n = W.lattice_2d(kw.get("L", 8), rng, pm_j=False)
    elif kind == "3d_gauss":
        i, j, K, n = W.lattice_3d(kw.get("L", 4), rng)
    elif kind == "3d_pm":
        i, j, K, n = W.lattice_3d(kw.get("L", 4), rng, pm_j=True)
    elif kind == "rr3":
        i, j, K, n = W.random_regular3(kw.get("n", 64), rng, gaussian=kw.get("gaussian", True),
                                       bias_sigma=kw.get("bias", 0.1))
    elif kind == "rr3_pm":
        i, j, K, n = W.random_regular3(kw.get("n", 64), rng, gaussian=False, bias_sigma=0.0)
    elif kind == "qubo":
        i, j, K, n = W.qubo_sparse(kw.get("n", 32), kw.get("density", 0.2), rng, qmax=kw.get("qmax", 100))
        qubo = True
    else:
        raise ValueError(kind)
    return O.Instance.from_lines(i, j, K, qubo=qubo), BqmIsingInstance.from_lines(i, j, K, qubo=qubo), (i, j, K)


def random_states(rng, count, n):
    return (rng.integers(0, 2, size=(count, n)) * 2 - 1).astype(np.int8)
