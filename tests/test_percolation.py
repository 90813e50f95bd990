"""Site percolation restatement (oracle/percolation.py, src/percolation.rs:11-75) pinned against
scipy's connected components and the closed forms of the occupancy rule; the percolate command's
arguments without a GPU."""
import os
import subprocess

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import oracle as O
from oracle import percolation as OP


def _square(L):
    """The reference test's lattice (src/percolation.rs:89-113): open boundaries, top row weight +1,
    bottom row -1."""
    i, j, K = [], [], []
    for r in range(L):
        for c in range(L):
            q = r * L + c
            if r < L - 1:
                i.append(q); j.append(q + L); K.append(1.0)
            if c < L - 1:
                i.append(q); j.append(q + 1); K.append(1.0)
            if r == 0:
                i.append(q); j.append(q); K.append(1.0)
            elif r == L - 1:
                i.append(q); j.append(q); K.append(-1.0)
    return np.array(i), np.array(j), np.array(K), L * L


def test_restatement_matches_connected_components():
    i, j, K, n = _square(8)
    inst = O.Instance.from_lines(i, j, K)
    rng = np.random.default_rng(3)
    for _ in range(40):
        occ = rng.random(n) < 0.6
        perc, gmax = OP.random_percolation(inst.row_ptr, inst.col, inst.bias, occ)
        rows = np.repeat(np.arange(n), np.diff(inst.row_ptr))
        keep = occ[rows] & occ[inst.col]
        g = sp.csr_matrix((np.ones(keep.sum()), (rows[keep], inst.col[keep])), shape=(n, n))
        nc, lab = connected_components(g, directed=False)
        sizes = np.bincount(lab[occ], minlength=nc) if occ.any() else np.zeros(1, int)
        assert gmax == (sizes.max() if occ.any() else 0)
        top = set(lab[k] for k in range(n) if occ[k] and inst.bias[k] > 0)
        bot = set(lab[k] for k in range(n) if occ[k] and inst.bias[k] < 0)
        assert perc == bool(top & bot)


def test_occupancy_rule_closed_forms():
    assert OP.occupancy(50, 1.0, 0, 7).all()
    assert not OP.occupancy(50, 0.0, 0, 7).any()
    occ = np.concatenate([OP.occupancy(400, 0.3, k, 11) for k in range(10)])
    assert abs(occ.mean() - 0.3) < 0.05
    # the draw is words 0, 1 of Philox keyed by the seed, counter {sample, site, 0, 4}
    w = O.philox([2, 5, 0, 4], [11, 0])
    v = (int(w[0]) << 32) | int(w[1])
    if v < 2 ** 64 - 8192:
        assert OP.occupancy(6, (v + 4096) / 18446744073709551616.0, 2, 11)[5]  # just above the draw
    if v > 8192:
        assert not OP.occupancy(6, (v - 4096) / 18446744073709551616.0, 2, 11)[5]  # just below


def test_percolate_command_line_arguments():
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "percolate")
    r = subprocess.run([cli, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "site percolation" in r.stdout
    r = subprocess.run([cli, "inst.txt"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--prob" in r.stderr
