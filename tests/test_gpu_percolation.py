"""Site percolation on the GPU (src/percolation.rs:11-75): per sample, the largest cluster and the
percolation flag equal the restatement in oracle/percolation.py on the same occupancy draws."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from oracle import percolation as OP
from tamc_amd import api, workloads as W
from tests.test_percolation import _square

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["square16", "rr3"])
def test_percolation_matches_restatement(gpu, case):
    if case == "square16":
        i, j, K, n = _square(16)
        p, samples, seed = 0.59, 120, 1234
    else:
        rng = np.random.default_rng(8)
        i, j, K, n = W.random_regular3(300, rng, gaussian=False, bias_sigma=0.0)
        w = rng.integers(-1, 2, n).astype(np.float64)
        i, j, K = np.concatenate([i, np.arange(n)]), np.concatenate([j, np.arange(n)]), np.concatenate([K, w])
        p, samples, seed = 0.45, 60, 99
    inst = api.BqmIsingInstance.from_lines(i, j, K)
    oi = O.Instance.from_lines(i, j, K)
    cl, pc = api.random_percolation(inst, p, samples, seed)
    for k in range(samples):
        occ = OP.occupancy(n, p, k, seed)
        perc, gmax = OP.random_percolation(oi.row_ptr, oi.col, oi.bias, occ)
        assert (int(cl[k]), bool(pc[k])) == (gmax, perc), k
    if case == "square16":
        # near p_c = 0.5927 of 2D site percolation a fair share of samples span the lattice
        assert 0.2 < pc.mean() < 0.9


def test_percolation_beyond_lds(gpu):
    # 30,000 sites: labels and counters in global scratch instead of LDS
    rng = np.random.default_rng(12)
    i, j, K, n = W.random_regular3(30000, rng, gaussian=False, bias_sigma=0.0)
    w = rng.integers(-1, 2, n).astype(np.float64)
    i, j, K = np.concatenate([i, np.arange(n)]), np.concatenate([j, np.arange(n)]), np.concatenate([K, w])
    p, samples, seed = 0.5, 6, 31
    inst = api.BqmIsingInstance.from_lines(i, j, K)
    oi = O.Instance.from_lines(i, j, K)
    cl, pc = api.random_percolation(inst, p, samples, seed)
    for k in range(samples):
        occ = OP.occupancy(n, p, k, seed)
        perc, gmax = OP.random_percolation(oi.row_ptr, oi.col, oi.bias, occ)
        assert (int(cl[k]), bool(pc[k])) == (gmax, perc), k


def test_percolation_edge_probabilities(gpu):
    i, j, K, n = _square(8)
    inst = api.BqmIsingInstance.from_lines(i, j, K)
    cl, pc = api.random_percolation(inst, 1.0, 4, 1)
    assert np.all(cl == n) and pc.all()
    cl, pc = api.random_percolation(inst, 0.0, 4, 1)
    assert np.all(cl == 0) and not pc.any()


def test_percolate_command_line(gpu):
    d = tempfile.mkdtemp()
    i, j, K, n = _square(16)
    W.write_instance_file(os.path.join(d, "g.txt"), i, j, K)
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "percolate")
    r = subprocess.run([cli, "-n", "50", "-p", "0.59", "-o", "out.txt", "g.txt"], capture_output=True, text=True,
                       cwd=d, timeout=120, env=dict(os.environ, TAMC_SEED="7"))
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[-3] == "# * Percolation results *" and lines[-2] == "samps,  num_perc,  frac_perc"
    inst = api.BqmIsingInstance.from_instance_file(os.path.join(d, "g.txt"))
    cl, pc = api.random_percolation(inst, 0.59, 50, 7)
    samps, num, frac = lines[-1].split(", ")
    assert int(samps) == 50 and int(num) == int(pc.sum()) and float(frac) == pc.sum() / 50
    assert open(os.path.join(d, "out.txt")).read() == "perc, max_cluster" + "".join(str(int(c)) for c in cl)
