"""Statistics of the production path (Philox streams, colour-ordered sweeps, all-pairs ICM).

Production runs use their own random streams and site order, so they are compared with the
reference's statistics rather than bit for bit:

* against exact enumeration of the Boltzmann distribution (N = 12), per beta: <E> and <q^2>;
* against the CPU restatement run with the reference's generator (Xoshiro256++, sequential site
  order, serial pt_loop), per beta: <E> and <q^2> on an 8x8 +-J lattice, and the ground state;
* at the benchmark configurations: the 16x16 +-J lattice of configuration 1 (all 32 betas) and
  the 12^3 Gaussian lattice of configuration 2 (the betas above its spin-glass transition), per
  beta: <E>, <q^2> and <chi^2> of a Gaussian suscept vector, against the reference generator;
* against the exact <E>(beta) of the 16x16 periodic ferromagnet (Kaufman's partition function).

Error bars come from binning.  Gate ("3 sigma"): with m comparisons in a table, each |z| must stay
below the two-sided family-wise 3-sigma bound Phi^-1(1 - 0.0027 / (2 m)) (Bonferroni), so the whole
table raises a false alarm with the probability of one 3-sigma test (0.27 %); per-comparison 3-sigma
gates would fail a correct sampler 23 % of the time at m = 96.  The mean of z^2 must also stay below
the chi-square 3-sigma bound 1 + 3 sqrt(2 / m).  The runs are seeded, so the tests are deterministic;
each prints its z-table.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
from scipy.stats import norm

import oracle as O
from tamc_amd import api, workloads as W
from tests.exact_ising import mean_energy

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


def binned_runs(series, bins_per_run=10):
    """Mean and error of the mean from equal bins of each run's post-warm-up series."""
    out = []
    for x in series:
        x = np.asarray(x, np.float64)
        out.append(x[: len(x) // bins_per_run * bins_per_run].reshape(bins_per_run, -1).mean(axis=1))
    b = np.concatenate(out)
    return b.mean(), b.std(ddof=1) / np.sqrt(len(b))


def gate(z, label):
    z = np.asarray(z, np.float64)
    crit = norm.isf(0.0027 / (2 * z.size))
    chi = 1 + 3 * np.sqrt(2.0 / z.size)
    print("%s: %d comparisons, max |z| %.2f (bound %.2f), mean z^2 %.2f (bound %.2f), |z| > 3: %d"
          % (label, z.size, np.abs(z).max(), crit, np.mean(z ** 2), chi, int((np.abs(z) > 3).sum())))
    assert np.all(np.abs(z) < crit), z
    assert np.mean(z ** 2) < chi, z


def _config_table(lines, B, bmin, bmax, C, T, runs, gated):
    """Production (GPU, one batched launch of `runs` seeds) against the reference generator (oracle,
    Xoshiro256++, `runs` seeds on host threads) on one instance, per beta: <E>, <q^2>, <chi^2>."""
    i, j, K, n = lines
    w = np.random.default_rng(3).standard_normal(n)[None]
    gi = api.BqmIsingInstance.from_lines(i, j, K).with_suscept(w)
    ps = [api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(bmin, bmax, B), lo_beta=1.0,
                          icm=True, num_replica_chains=C, sample=1, sample_states=None, seed=1000 + k)
          for k in range(runs)]
    betas, lo = ps[0].resolve()
    gpu_runs, _ = api.run_pt_batch([gi] * runs, ps)

    def ref(k):
        oi = O.Instance.from_lines(i, j, K)
        oi.set_suscept(w)
        return O.pt_run(oi, betas, T, num_chains=C, warmup_fraction=0.5, lo_beta_idx=lo, icm=True, sample=1,
                        rng_mode=O.RNG_XOSHIRO, seed=2000 + k)
    with ThreadPoolExecutor(max_workers=16) as ex:
        ref_runs = list(ex.map(ref, range(runs)))
    z, rows = [], []
    for b in range(B):
        row = []
        for name, f in (("E", lambda r: r["e"][b]), ("q2", lambda r: r["q"][b].astype(np.float64) ** 2),
                        ("chi2", lambda r: r["suscept"][b][0].astype(np.float64) ** 2)):
            (m1, s1), (m2, s2) = binned_runs([f(r) for r in gpu_runs]), binned_runs([f(r) for r in ref_runs])
            s = np.hypot(s1, s2)
            zz = (m1 - m2) / s if s > 0 else (0.0 if m1 == m2 else np.inf)
            row.append(zz)
            if gated(betas[b]):
                z.append(zz)
        rows.append(row)
        print("beta %.4f  z(E) %+.2f  z(q^2) %+.2f  z(chi^2) %+.2f%s"
              % (betas[b], row[0], row[1], row[2], "" if gated(betas[b]) else "  (not gated)"))
    gs_gpu = min(r["gs_energies"][-1] for r in gpu_runs)
    gs_ref = min(r["gs_energies"][-1] for r in ref_runs)
    return z, gs_gpu, gs_ref


def test_config1_statistics_vs_reference_generator(gpu):
    lines = W.lattice_2d(16, np.random.default_rng(101), pm_j=True)
    z, gs_gpu, gs_ref = _config_table(lines, 32, 0.2, 5.0, 2, 10000, 8, lambda b: True)
    gate(z, "configuration 1")
    assert gs_gpu == gs_ref  # the same ground-state energy (exact for +-J)


def test_config2_statistics_vs_reference_generator(gpu):
    # 3D 12^3 Gaussian glass: the slots above the spin-glass transition (beta < 0.8; T_c ~ 0.95)
    # equilibrate within these runs; below it the two dynamics relax at different rates within
    # 2000 sweeps, so those slots are printed but not gated
    lines = W.lattice_3d(12, np.random.default_rng(102), pm_j=False)
    z, _, _ = _config_table(lines, 48, 0.1, 5.0, 4, 2000, 8, lambda b: b < 0.8)
    gate(z, "configuration 2, beta < 0.8")


def test_ferromagnet_16x16_matches_exact_energy(gpu):
    L, B, C, T, runs = 16, 32, 2, 10000, 8
    i, j, K, n = W.lattice_2d(L, ferro=True)
    gi = api.BqmIsingInstance.from_lines(i, j, K)
    ps = [api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, B), lo_beta=1.0,
                          icm=True, num_replica_chains=C, sample=1, sample_states=None, seed=300 + k)
          for k in range(runs)]
    betas, _ = ps[0].resolve()
    res, _ = api.run_pt_batch([gi] * runs, ps)
    z = []
    for b in range(B):
        exact = mean_energy(float(betas[b]), L, L)
        m, s = binned_runs([r["e"][b] for r in res])
        if s == 0:
            assert abs(m - exact) < 1e-3, (betas[b], m, exact)  # frozen in the ground state
            zz = 0.0
        else:
            zz = (m - exact) / s
        z.append(zz)
        print("beta %.4f  <E> %.4f +- %.4f  exact %.4f  z %+.2f" % (betas[b], m, s, exact, zz))
    gate(z, "ferromagnet 16x16 vs exact")
