"""GPU parity at the shapes of the five benchmark configurations, against the CPU restatement
(oracle/) on the same inputs and Philox keys.

  c1  2D 16x16 periodic +-J, 32 geometric betas 0.2-5.0, 2 chains, 2000 sweeps (full run), ICM from
      beta 1.0, sample 32/32, sample_limiting 2, two suscept vectors: every output bit-exact.
  c2  3D 12^3 periodic J~N(0,1), 48 betas 0.1-5.0, 4 chains, 150 of the 10000 sweeps, two suscept
      vectors: every output bit-exact (the oracle's production mode restates the device's exact
      tracked-energy accounting).
  c3  the benchmark's own 256 random 3-regular graphs (N ~ U{200..4000}, J~N(0,1), h~N(0,0.1)) in
      one autotuned launch and in 16-wave launches with either task tiling and either ICM growth
      order, 16 sweeps with ICM from sweep 0: every instance against the oracle.
  c4  64 QUBOs N=2048 at 5 % (Q ~ U{-100..100}, diagonal included), 1024 restarts, one batched SA
      launch over 8 sweeps of a linear 0.1 -> 10 schedule: three instances state by state.
  c5  a seeded mini-stream from the serving distribution plus its largest shapes (2D L=128,
      3D L=24, random 3-regular N=16384), served together by the continuous-batching server:
      each result equals its single run and the oracle bit for bit.

The instances are those of bench.py (same generators and seeds).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from tamc_amd import api
from tamc_amd import workloads as W
from tests.test_gpu_parity import EXACT_KEYS, assert_same

pytestmark = pytest.mark.gpu


def _suscept_vectors(rng, n, coords=None):
    """Two coefficient vectors: a Gaussian one (its f64 sum depends on the site order) and a
    staggered one built from site coordinates."""
    w0 = rng.standard_normal(n)
    if coords is None:
        w1 = np.where(np.arange(n) % 2 == 0, 1.0, -1.0) / n
    else:
        w1 = np.where(coords.sum(axis=1) % 2 == 0, 1.0, -1.0) * (1.0 + 0.01 * coords[:, 0])
    return np.stack([w0, w1])


def _oracle(oi, p, betas, lo, init=None):
    return O.pt_run(oi, betas, p.num_sweeps, num_chains=p.num_replica_chains, warmup_fraction=p.warmup_fraction,
                    lo_beta_idx=lo, icm=p.icm, threaded=p.threads > 1, sample=p.sample,
                    sample_states=p.sample_states, sample_limiting=p.sample_limiting or 0, rng_mode=O.RNG_PHILOX,
                    seed=p.seed, sweep_order=p.sweep_order, init_states=init)


def _pair(lines, qubo=False):
    i, j, K = lines[:3]
    return O.Instance.from_lines(i, j, K, qubo=qubo), api.BqmIsingInstance.from_lines(i, j, K, qubo=qubo)


def test_c1_full_run_with_suscepts(gpu):
    L = 16
    lines = W.lattice_2d(L, np.random.default_rng(101), pm_j=True)
    oi, gi = _pair(lines)
    yx = np.stack(np.divmod(np.arange(L * L), L), axis=1)
    w = _suscept_vectors(np.random.default_rng(1), L * L, yx)
    gi.with_suscept(w)
    oi.set_suscept(w)
    p = api.PtIcmParams(num_sweeps=2000, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 32),
                        lo_beta=1.0, icm=True, num_replica_chains=2, threads=1, sample=32, sample_states=32,
                        sample_limiting=2, seed=101)
    runner = api.PtIcmRunner(gi, p)
    assert runner.lo_beta_idx == 16
    g = runner.run()
    o = _oracle(oi, p, runner.betas, runner.lo_beta_idx)
    assert o["error"] == 0
    assert g["n_meas"] == 33 and g["suscept"].shape == (32, 2, 33)
    assert_same(g, o, exact_energy=True)
    # the suscept of the uniform-sign vector is q itself up to the coefficient scale
    assert np.any(g["suscept"] != 0)


@pytest.mark.parametrize("threads", [1, 4])
def test_c2_shape_with_suscepts(gpu, threads):
    L = 12
    lines = W.lattice_3d(L, np.random.default_rng(102), pm_j=False)
    oi, gi = _pair(lines)
    zyx = np.stack(np.unravel_index(np.arange(L ** 3), (L, L, L)), axis=1)
    w = _suscept_vectors(np.random.default_rng(2), L ** 3, zyx)
    gi.with_suscept(w)
    oi.set_suscept(w)
    p = api.PtIcmParams(num_sweeps=150, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.1, 5.0, 48),
                        lo_beta=1.0, icm=True, num_replica_chains=4, threads=threads, sample=8, sample_states=16,
                        sample_limiting=1, seed=102 + threads)
    runner = api.PtIcmRunner(gi, p)
    g = runner.run()
    o = _oracle(oi, p, runner.betas, runner.lo_beta_idx)
    assert o["error"] == 0
    assert g["suscept"].shape == (48, 2, g["n_meas"] * 2)
    assert_same(g, o, exact_energy=True)


def _c3_workload(T):
    ws = W.config3_instances(count=256, seed=103)
    insts = [api.BqmIsingInstance.from_lines(i, j, K) for (i, j, K, n) in ws]
    params = [api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 32),
                              lo_beta=1.0, icm=True, num_replica_chains=2, threads=1, sample=4, sample_states=4,
                              sample_limiting=2, seed=1000 + k) for k in range(len(ws))]
    return ws, insts, params


def _oracle_many(ws, params, betas, lo, qubo=False):
    def one(k):
        i, j, K = ws[k][:3]
        return _oracle(O.Instance.from_lines(i, j, K, qubo=qubo), params[k], betas, lo)
    with ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(one, range(len(ws))))


def test_c3_bench_batch_autotuned_and_16_waves(gpu):
    T = 16
    ws, insts, params = _c3_workload(T)
    betas, lo = params[0].resolve()
    tuned, tst = api.run_pt_batch(insts, params, autotune=True)
    assert tst["waves_per_group"] in (4, 8, 16) and tst["launches"] in (1, 4) and tst["groups"] == 256
    assert tst["task_tiling"] in (1, 2)
    wide, wst = api.run_pt_batch(insts, params, waves_per_group=16, task_tiling=1, icm_growth=1)
    assert wst["waves_per_group"] == 16
    runs, _ = api.run_pt_batch(insts, params, waves_per_group=16, task_tiling=2, icm_growth=2)
    assert max(x.n for x in insts) > 3900  # the largest group runs the LDS-resident growth path
    oracles = _oracle_many(ws, params, betas, lo)
    for k, o in enumerate(oracles):
        assert o["error"] == 0
        assert o["icm_moves"] > 0
        assert_same(tuned[k], o, exact_energy=True, ctx="instance %d" % k)
        for key in EXACT_KEYS + ["e", "gs_energies", "final_energies"]:
            assert np.array_equal(wide[k][key], tuned[k][key]), (k, key)
            assert np.array_equal(runs[k][key], tuned[k][key]), (k, key)


def test_c4_sa_batch_shape(gpu):
    betas = (0.1 + 9.9 * np.arange(8) / 7).astype(np.float32)
    lines, insts = [], []
    for k in range(64):
        q = W.qubo_sparse(2048, 0.05, np.random.default_rng(104 + k), qmax=100)
        lines.append(q)
        insts.append(api.BqmIsingInstance.from_lines(q[0], q[1], q[2], qubo=True))
    threads = {63: 3}  # the reference keeps threads * floor(R / threads) restarts (src/sa.rs:147,161)
    ps = [api.SaParams(beta=api.BetaOptions.Arr(betas), num_replicas=1024, threads=threads.get(k, 1), seed=k)
          for k in range(64)]
    res, st = api.run_sa_batch(insts, ps)
    assert st["attempts"] == sum(8 * 2048 * (1023 if k == 63 else 1024) for k in range(64))

    def one(k):
        q = lines[k]
        return O.sa_run(O.Instance.from_lines(q[0], q[1], q[2], qubo=True), betas, 1024,
                        threads=threads.get(k, 1), rng_mode=O.RNG_PHILOX, seed=k)
    with ThreadPoolExecutor(max_workers=3) as ex:
        oracles = dict(zip([0, 31, 63], ex.map(one, [0, 31, 63])))
    for k, o in oracles.items():
        g = res[k]
        assert g["R"] == o["R"] == (1023 if k == 63 else 1024)
        assert np.array_equal(g["final_states"], o["final_states"]), k
        assert np.array_equal(g["energies"], o["energies"]), k
        assert g["min_energy"] == o["min_energy"]


def _c5_requests():
    reqs = [(kind, size, lines) for (_, kind, size, lines) in W.serving_requests(12, 200.0, seed=105)]
    reqs.append(("2d", 128, W.lattice_2d(128, np.random.default_rng(8), pm_j=True)))
    reqs.append(("3d", 24, W.lattice_3d(24, np.random.default_rng(7), pm_j=False)))
    reqs.append(("rr3", 16384, W.random_regular3(16384, np.random.default_rng(5), gaussian=True)))
    return reqs


def test_c5_mini_stream_served(gpu):
    reqs = _c5_requests()
    T = 6
    insts, oinsts, params = [], [], []
    for k, (kind, size, lines) in enumerate(reqs):
        oi, gi = _pair(lines)
        if k % 4 == 1:  # some requests carry suscept vectors
            w = _suscept_vectors(np.random.default_rng(50 + k), gi.n)
            gi.with_suscept(w)
            oi.set_suscept(w)
        insts.append(gi)
        oinsts.append(oi)
        params.append(api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5,
                                      beta=api.BetaOptions.Geometric(0.2, 5.0, 32), lo_beta=1.0, icm=True,
                                      num_replica_chains=2, sample=1, sample_states=2, sample_limiting=2,
                                      seed=5000 + k))
    with api.Server(slice_ms=1.0) as srv:
        tickets = [srv.submit(g, p) for g, p in zip(insts, params)]
        # a resident instance is immutable: its suscept vectors cannot be replaced meanwhile
        with pytest.raises(api._lib.TamcError):
            insts[1].with_suscept(np.zeros((1, insts[1].n)))
        served = [srv.wait(t)[0] for t in tickets]
    insts[1].with_suscept(_suscept_vectors(np.random.default_rng(51), insts[1].n))  # allowed again
    singles = [api.PtIcmRunner(g, p).run() for g, p in zip(insts, params)]
    betas, lo = params[0].resolve()

    def one(k):
        return _oracle(oinsts[k], params[k], betas, lo)
    with ThreadPoolExecutor(max_workers=16) as ex:
        oracles = list(ex.map(one, range(len(reqs))))
    for k in range(len(reqs)):
        for key in EXACT_KEYS + ["e", "gs_energies", "final_energies", "suscept"]:
            assert np.array_equal(served[k][key], singles[k][key]), (k, key)
        assert oracles[k]["error"] == 0
        assert_same(served[k], oracles[k], exact_energy=True, ctx="request %d" % k)
