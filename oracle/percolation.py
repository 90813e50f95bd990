"""Restatement of random_percolation (src/percolation.rs:11-75) with the build's occupancy draws.

Test infrastructure: the product package never imports this module.

Occupancy of site i in sample k: rand 0.8's Bernoulli rule (v < p * 2^64, always for p = 1) on the
64-bit draw (w0 << 32 | w1) of Philox4x32-10 with key = seed, counter {k, i, 0, 4}.  The largest
cluster is the largest strongly connected component of the occupied subgraph (kosaraju_scc on an
undirected graph = connected components); percolation is the reference's depth-first search from
every occupied negative-weight site, stopped at the first tree edge into an occupied positive-weight
site.
"""
import numpy as np

from .oracle import philox


def occupancy(n, p, k, seed):
    if p >= 1.0:
        return np.ones(n, bool)
    p_int = int(p * 18446744073709551616.0)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    occ = np.zeros(n, bool)
    for i in range(n):
        w = philox([k, i, 0, 4], key)
        occ[i] = ((int(w[0]) << 32) | int(w[1])) < p_int
    return occ


def random_percolation(row_ptr, col, weight, occ):
    """(path found, largest cluster size) for one occupancy vector."""
    n = len(occ)
    comp = -np.ones(n, np.int64)
    sizes = []
    for s in range(n):
        if not occ[s] or comp[s] >= 0:
            continue
        stack, comp[s], size = [s], len(sizes), 0
        while stack:
            u = stack.pop()
            size += 1
            for e in range(row_ptr[u], row_ptr[u + 1]):
                v = int(col[e])
                if occ[v] and comp[v] < 0:
                    comp[v] = comp[s]
                    stack.append(v)
        sizes.append(size)
    g_max = max(sizes) if sizes else 0
    init = [i for i in range(n) if occ[i] and weight[i] < 0]
    tgt = set(i for i in range(n) if occ[i] and weight[i] > 0)
    if not init or not tgt:
        return False, g_max
    # depth-first search from all init sites (petgraph depth_first_search with several starts)
    seen = np.zeros(n, bool)
    for s in init:
        if seen[s]:
            continue
        seen[s] = True
        stack = [s]
        while stack:
            u = stack.pop()
            for e in range(row_ptr[u], row_ptr[u + 1]):
                v = int(col[e])
                if occ[v] and not seen[v]:
                    if v in tgt:
                        return True, g_max
                    seen[v] = True
                    stack.append(v)
    return False, g_max
