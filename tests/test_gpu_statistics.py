"""Statistics of the production path (Philox streams, level-ordered sweeps, all-pairs ICM).

Production runs use their own random streams, so they are compared with the reference's
statistics rather than bit for bit:

* against exact enumeration of the Boltzmann distribution (N = 12), per beta: <E> and <q^2>;
* against the CPU restatement run with the reference's generator (Xoshiro256++, serial pt_loop),
  per beta: <E> and <q^2> on an 8x8 +-J lattice, and the ground-state energy.

Error bars come from binning (100 bins of consecutive measurements).  The runs are seeded, so
the test is deterministic; the gate is |z| < 4 for every comparison, and the mean of z^2 stays
below 2.5 (both chi-square style checks would flag a biased sampler at these run lengths).
"""
import numpy as np
import pytest

import oracle as O
from tamc_amd import api, workloads as W

pytestmark = pytest.mark.gpu


def binned(x, nbins=100):
    x = np.asarray(x, np.float64)
    x = x[: len(x) // nbins * nbins].reshape(nbins, -1).mean(axis=1)
    return x.mean(), x.std(ddof=1) / np.sqrt(nbins)


def exact_moments(inst, n, beta):
    states = (((np.arange(2 ** n)[:, None] >> np.arange(n)) & 1) * -2 + 1).astype(np.int8)
    E = np.array([inst.energy(s) for s in states], np.float64)
    w = np.exp(-beta * (E - E.min()))
    w /= w.sum()
    m2 = (w[:, None, None] * (states[:, :, None] * states[:, None, :])).sum(axis=0)  # <s_i s_j>
    return (w * E).sum(), (m2 ** 2).sum()  # <q^2> of two independent replicas = sum_ij <s_i s_j>^2


def per_beta(res, key, B, width):
    arr = np.asarray(res[key]).reshape(B, -1, width)
    return [arr[b].reshape(-1) for b in range(B)]


def test_production_matches_enumeration(gpu):
    rng = np.random.default_rng(21)
    i, j, K, n = W.random_regular3(12, rng, gaussian=True, bias_sigma=0.3)
    gi = api.BqmIsingInstance.from_lines(i, j, K)
    oi = O.Instance.from_lines(i, j, K)
    B, C, T = 6, 2, 40000
    p = api.PtIcmParams(num_sweeps=T, warmup_fraction=0.05, beta=api.BetaOptions.Geometric(0.3, 2.5, B), lo_num_beta=3,
                        num_replica_chains=C, sample=1, seed=2024)
    res = api.PtIcmRunner(gi, p).run()
    betas = p.resolve()[0]
    es = per_beta(res, "e", B, C)
    qs = per_beta(res, "q", B, 1)
    z = []
    for b in range(B):
        eE, eQ2 = exact_moments(oi, n, float(betas[b]))
        m, s = binned(es[b])
        z.append((m - eE) / s)
        m, s = binned(np.asarray(qs[b], np.float64) ** 2)
        z.append((m - eQ2) / s)
    z = np.array(z)
    assert np.all(np.abs(z) < 4), z
    assert np.mean(z ** 2) < 2.5, z


def test_production_matches_reference_generator(gpu):
    rng = np.random.default_rng(22)
    i, j, K, n = W.lattice_2d(8, rng, pm_j=True)
    gi = api.BqmIsingInstance.from_lines(i, j, K)
    oi = O.Instance.from_lines(i, j, K)
    B, C, T = 12, 2, 20000
    p = api.PtIcmParams(num_sweeps=T, warmup_fraction=0.1, beta=api.BetaOptions.Geometric(0.2, 3.0, B), lo_beta=1.0,
                        num_replica_chains=C, sample=1, seed=77)
    betas, lo = p.resolve()
    g = api.PtIcmRunner(gi, p).run()
    o = O.pt_run(oi, betas, T, num_chains=C, warmup_fraction=0.1, lo_beta_idx=lo, icm=True, sample=1,
                 rng_mode=O.RNG_XOSHIRO, seed=5)
    z = []
    for key, width in (("e", C), ("q", 1)):
        ga, oa = per_beta(g, key, B, width), per_beta(o, key, B, width)
        for b in range(B):
            x, y = np.asarray(ga[b], np.float64), np.asarray(oa[b], np.float64)
            if key == "q":
                x, y = x ** 2, y ** 2
            (m1, s1), (m2, s2) = binned(x), binned(y)
            z.append((m1 - m2) / np.hypot(s1, s2))
    z = np.array(z)
    assert np.all(np.abs(z) < 4), z
    assert np.mean(z ** 2) < 2.5, z
    # the same ground-state energy (exact in fp32 for +-J)
    assert g["gs_energies"][-1] == o["gs_energies"][-1]
