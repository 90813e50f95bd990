"""GPU parity: the HIP path through the C ABI against the CPU restatement (oracle/) on the same inputs.

Every output is compared bit-exactly.  The oracle's production mode (Philox keys, colour-class
order) also restates the production path's tracked-energy accounting -- each sweep's accepted dE
summed exactly in 2^-32 fixed point, and an ICM move adding the cluster's exact energy change --
so tracked energies, and with them every exchange decision, are bit-exact for Gaussian couplings
too.  (With the reference's f32 accumulation instead, Gaussian energies differ in the last bits,
and an exchange decision within an ulp of its threshold can then differ.)
"""
import numpy as np
import pytest

import oracle as O
from tamc_amd import api
from tests.helpers import make_pair, random_states

pytestmark = pytest.mark.gpu

EXACT_KEYS = ["q", "samples", "gs_steps", "gs_states", "acceptance_counts", "final_states", "final_ids",
              "num_acceptances", "round_trips", "tags", "diffusion_hist"]


def assert_same(g, o, exact_energy=True, tol=2e-4, ctx=""):
    """Exact for +-J/QUBO.  For Gaussian couplings the ground-state trace records strict f32
    improvements, so re-finding a state whose tracked energy differs in the last bits adds records:
    there the trace is compared by its final energy (within tol) and every other output exactly."""
    assert g["n_meas"] == o["n_meas"]
    keys = EXACT_KEYS if exact_energy else [k for k in EXACT_KEYS if k not in ("gs_steps", "gs_states")]
    for k in keys:
        assert g[k].shape == o[k].shape, (ctx, k)
        assert np.array_equal(g[k], o[k]), (ctx, k)
    assert g["icm_moves"] == o["icm_moves"], ctx
    if exact_energy:
        for k in ["e", "gs_energies", "final_energies", "suscept"]:
            assert g[k].shape == o[k].shape, (ctx, k)
            assert np.array_equal(g[k], o[k]), (ctx, k)
    else:
        for k in ["e", "final_energies", "suscept"]:
            np.testing.assert_allclose(g[k], o[k], rtol=tol, atol=tol, err_msg=k)
        np.testing.assert_allclose(g["gs_energies"][-1], o["gs_energies"][-1], rtol=tol, atol=tol)


def test_energy_delta_overlap_pack(gpu):
    rng = np.random.default_rng(1)
    for kind, kw in [("2d_pm", {}), ("3d_gauss", {}), ("rr3", dict(n=50)), ("qubo", dict(n=37)),
                     ("2d_gauss", dict(L=5))]:
        oi, gi, _ = make_pair(kind, seed=3, **kw)
        S = random_states(rng, 17, oi.n)
        ge = gi.energy(S)
        oe = np.array([oi.energy(s) for s in S], np.float32)
        assert np.array_equal(ge, oe), kind
        gd = gi.delta_energy(S)
        od = np.array([[oi.delta_energy(s, i) for i in range(oi.n)] for s in S], np.float32)
        assert np.array_equal(gd, od), kind
    S1, S2 = random_states(rng, 9, 77), random_states(rng, 9, 77)
    assert np.array_equal(api.overlap(S1, S2), [O.overlap(a, b) for a, b in zip(S1, S2)])
    assert np.array_equal(api.pack_bytes(S1), np.array([O.pack_bytes(s) for s in S1]))


@pytest.mark.parametrize("kind,kw", [("2d_pm", {}), ("rr3", dict(n=96)), ("3d_pm", dict(L=3)),
                                     ("qubo", dict(n=40, density=0.1))])
def test_houdayer_flip_sets(gpu, kind, kw):
    oi, gi, _ = make_pair(kind, seed=7, **kw)
    rng = np.random.default_rng(11)
    npairs = 40  # more than one 32-pair chunk
    S1, S2 = random_states(rng, npairs, oi.n), random_states(rng, npairs, oi.n)
    S2[0] = S1[0]  # a pair without q<0 sites: no move
    ks = []
    for a, b in zip(S1, S2):
        cnt = O.num_negative_overlap(a, b)
        ks.append(int(rng.integers(0, cnt)) if cnt else 0)
    g1, g2, sizes, cl = gi.houdayer(S1, S2, ks)
    for p in range(npairs):
        a, b = S1[p].copy(), S2[p].copy()
        r, ocl = oi.houdayer_k(a, b, ks[p])
        assert sizes[p] == max(r, 0)
        assert np.array_equal(g1[p], a) and np.array_equal(g2[p], b)
        assert np.array_equal(cl[p], ocl)


@pytest.mark.parametrize("long_pairs", [(3, 17, 33), tuple(range(20)) + (35,)])
def test_houdayer_long_clusters(gpu, long_pairs):
    # q<0 on a serpentine through a periodic 32x32 lattice: the even rows but the last, joined
    # through one site of each odd row, alternately at columns 0 and 16, so that the cluster is
    # 527 sites whose far end is 150 to 270 hops from the seed.  With at most 16 such pairs among
    # 32 the clusters are finished by component labelling after the first 10 passes; with 20 the
    # bit-parallel growth goes on until no more than 16 are still growing.  Both must give the
    # reference's flip sets.
    L = 32
    oi, gi, _ = make_pair("2d_pm", seed=9, L=L)
    rng = np.random.default_rng(13)
    npairs = 40
    S1 = random_states(rng, npairs, oi.n)
    S2 = random_states(rng, npairs, oi.n)
    r, c = np.divmod(np.arange(oi.n), L)
    snake = (r < L - 1) & ((r % 2 == 0) | (c == np.where((r // 2) % 2 == 0, 0, L // 2)))
    for p in long_pairs:
        S2[p] = np.where(snake, -S1[p], S1[p])
    ks = []
    for a, b in zip(S1, S2):
        cnt = O.num_negative_overlap(a, b)
        ks.append(int(rng.integers(0, cnt)) if cnt else 0)
    g1, g2, sizes, cl = gi.houdayer(S1, S2, ks)
    for p in range(npairs):
        a, b = S1[p].copy(), S2[p].copy()
        rr, ocl = oi.houdayer_k(a, b, ks[p])
        assert sizes[p] == max(rr, 0)
        assert np.array_equal(g1[p], a) and np.array_equal(g2[p], b)
        assert np.array_equal(cl[p], ocl)
    for p in long_pairs:
        assert sizes[p] == snake.sum()


def _run_both(kind, kw, B=8, C=2, T=60, lo=4, threads=1, seed=5, icm=True, sample=4, sstates=4, limiting=0,
              bmin=0.2, bmax=5.0, order="colour"):
    oi, gi, _ = make_pair(kind, seed=seed, **kw)
    params = api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(bmin, bmax, B),
                             lo_num_beta=B - lo, icm=icm, num_replica_chains=C, threads=threads, sample=sample,
                             sample_states=sstates, sample_limiting=limiting, seed=seed * 1000 + 17,
                             sweep_order=order)
    runner = api.PtIcmRunner(gi, params)
    g = runner.run()
    o = O.pt_run(oi, runner.betas, T, num_chains=C, warmup_fraction=0.5, lo_beta_idx=runner.lo_beta_idx, icm=icm,
                 threaded=threads > 1, sample=sample, sample_states=sstates, sample_limiting=limiting,
                 rng_mode=O.RNG_PHILOX, seed=params.seed, sweep_order=order)
    return g, o


@pytest.mark.parametrize("kind,kw,C,threads,limiting", [
    ("2d_pm", {}, 2, 1, 0), ("2d_pm", {}, 2, 4, 1), ("2d_ferro", dict(L=6), 4, 1, 2), ("3d_pm", dict(L=3), 3, 1, 0),
    ("rr3_pm", dict(n=64), 2, 1, 0), ("qubo", dict(n=24, density=0.2), 2, 1, 0), ("2d_pm", dict(L=5), 1, 1, 0)])
def test_pt_production_exact(gpu, kind, kw, C, threads, limiting):
    g, o = _run_both(kind, kw, C=C, threads=threads, limiting=limiting)
    assert o["error"] == 0
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("kind,kw,C,threads", [("2d_pm", {}, 2, 1), ("rr3_pm", dict(n=64), 2, 4),
                                               ("qubo", dict(n=24, density=0.2), 3, 1), ("3d_pm", dict(L=3), 2, 1)])
def test_pt_production_sequential_order(gpu, kind, kw, C, threads):
    # level schedule: the reference's own sequential site order
    g, o = _run_both(kind, kw, C=C, threads=threads, order="sequential")
    assert o["error"] == 0
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("kind,kw", [("rr3", dict(n=120)), ("3d_gauss", dict(L=4)), ("2d_gauss", dict(L=6))])
def test_pt_production_gaussian(gpu, kind, kw):
    g, o = _run_both(kind, kw, T=30, B=6, lo=2)
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("kind,kw", [("2d_pm", dict(L=128)), ("3d_pm", dict(L=24)), ("rr3", dict(n=16384))])
def test_pt_production_large_instances(gpu, kind, kw):
    # N up to 16384 with 32 betas x 2 chains: the spins still fit in LDS, the ICM cluster words and
    # the neighbour table do not, so ICM runs on its global-memory scratch
    g, o = _run_both(kind, kw, B=32, lo=16, T=6, sample=2, sstates=3, limiting=2)
    assert o["error"] == 0
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("kind,kw,B,C,T", [("2d_pm", dict(L=182), 32, 2, 6), ("rr3", dict(n=70000), 32, 2, 4),
                                           ("3d_gauss", dict(L=20), 96, 2, 4), ("2d_pm", dict(L=100), 96, 3, 4)])
def test_pt_production_spins_in_global_memory(gpu, kind, kw, B, C, T):
    # chain pairs whose spin planes do not fit in LDS: 33,124 sites at 64 replicas (265 KB of
    # planes), 70,000 sites (beyond the 16-bit neighbour lists, so ICM walks the coupling CSR), and
    # 8,000 and 10,000 sites at 192 replicas (three plane words per site; with three chains the
    # third runs alone).  The planes stay in the group's global state array and the clusters grow
    # in global scratch.
    g, o = _run_both(kind, kw, B=B, C=C, T=T, lo=B // 2, sample=2, sstates=3, limiting=2)
    assert o["error"] == 0
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("kind,kw,B,C,threads", [("2d_pm", dict(L=8), 200, 2, 1), ("rr3", dict(n=100), 300, 3, 4),
                                                   ("2d_gauss", dict(L=6), 512, 2, 1)])
def test_pt_production_long_ladders(gpu, kind, kw, B, C, threads):
    # ladders of 200 to 512 betas: chain pairs of 400 to 1024 replicas (7 to 16 plane words per
    # site), exchange walked by one lane, ICM over several 32-pair chunks
    g, o = _run_both(kind, kw, B=B, C=C, T=12, lo=B // 3, threads=threads, sample=3, sstates=4, limiting=1)
    assert o["error"] == 0
    assert_same(g, o, exact_energy=True)


def test_icm_growth_every_sites_per_thread(gpu):
    # one wave per work group (64 threads): the cluster growth keeps 4, 8 or 16 sites per thread in
    # registers, for neighbour rows of one 64-bit word (width 4) and of two (3D lattices, width 6);
    # beyond 16 sites per thread it loops over memory
    cases = [("2d_pm", dict(L=12)), ("rr3_pm", dict(n=400)), ("2d_pm", dict(L=30)),
             ("3d_pm", dict(L=6)), ("3d_pm", dict(L=7)), ("3d_pm", dict(L=10)),
             ("2d_pm", dict(L=34)), ("3d_pm", dict(L=11))]
    B, C, T = 8, 2, 12
    insts, params, oracles = [], [], []
    for k, (kind, kw) in enumerate(cases):
        oi, gi, _ = make_pair(kind, seed=40 + k, **kw)
        insts.append(gi)
        oracles.append(oi)
        params.append(api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, B),
                                      lo_num_beta=4, icm=True, num_replica_chains=C, threads=1, sample=2,
                                      sample_states=3, sample_limiting=2, seed=900 + k))
    assert [g.n for g in insts] == [144, 400, 900, 216, 343, 1000, 1156, 1331]
    # strided sites per thread, and compact 64-site tiles closed one wave at a time
    strided, st = api.run_pt_batch(insts, params, waves_per_group=1, icm_growth=1)
    assert st["waves_per_group"] == 1 and st["icm_growth"] == 1
    tiles, st2 = api.run_pt_batch(insts, params, waves_per_group=1, icm_growth=2)
    assert st2["icm_growth"] == 2
    for k, (oi, p) in enumerate(zip(oracles, params)):
        runner = api.PtIcmRunner(insts[k], p)
        o = O.pt_run(oi, runner.betas, T, num_chains=C, warmup_fraction=0.5, lo_beta_idx=runner.lo_beta_idx, icm=True,
                     threaded=False, sample=2, sample_states=3, sample_limiting=2, rng_mode=O.RNG_PHILOX, seed=p.seed)
        assert o["error"] == 0
        assert_same(strided[k], o, exact_energy=True)
        assert_same(tiles[k], o, exact_energy=True)


def test_icm_staged_growth_over_memory(gpu):
    # 2D L=100 with 32 betas x 2 chains: the ICM words do not fit in LDS beside the spins, so the
    # growth runs in the spin planes' LDS while the planes are parked; at 8 waves (512 threads)
    # 10000 sites exceed 16 per thread, so it loops over memory there
    oi, gi, _ = make_pair("2d_pm", seed=61, L=100)
    p = api.PtIcmParams(num_sweeps=6, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 32),
                        lo_num_beta=16, icm=True, num_replica_chains=2, threads=1, sample=2, sample_states=3,
                        sample_limiting=2, seed=4242)
    batch, st = api.run_pt_batch([gi], [p], waves_per_group=8)
    assert st["waves_per_group"] == 8
    runner = api.PtIcmRunner(gi, p)
    o = O.pt_run(oi, runner.betas, 6, num_chains=2, warmup_fraction=0.5, lo_beta_idx=runner.lo_beta_idx, icm=True,
                 threaded=False, sample=2, sample_states=3, sample_limiting=2, rng_mode=O.RNG_PHILOX, seed=p.seed)
    assert o["error"] == 0
    assert_same(batch[0], o, exact_energy=True)


def test_pt_large_ladder_two_words(gpu):
    # 48 betas x 2 chains = 96 replicas: two bit-plane words per site
    g, o = _run_both("3d_pm", dict(L=3), B=48, lo=28, T=24, sample=3, sstates=5)
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("threads", [1, 2])
def test_pt_ladder_longer_than_a_wave(gpu, threads):
    # 80 betas: the exchange walk no longer fits one lane per step of a wave; 160 replicas, 3 words
    g, o = _run_both("2d_pm", dict(L=6), B=80, lo=60, T=20, threads=threads, sample=3, sstates=5)
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("kind,kw,B,lo,C,threads,T", [
    ("3d_pm", dict(L=3), 48, 4, 2, 1, 16),        # 44 ICM slots: chunks of 32 and 12
    ("2d_pm", dict(L=6), 80, 2, 4, 2, 12),        # 78 slots in three chunks, two chain pairs
    ("rr3", dict(n=200), 64, 0, 2, 1, 8),         # every slot, Gaussian couplings
    ("2d_pm", dict(L=88), 40, 0, 2, 1, 3)])       # ICM words in global memory
def test_icm_more_than_32_slots(gpu, kind, kw, B, lo, C, threads, T):
    # a cluster word holds 32 pairs: the kernel runs the pairs of a chain pair 32 slots at a time
    g, o = _run_both(kind, kw, B=B, lo=lo, C=C, threads=threads, T=T, sample=2, sstates=3, limiting=2)
    assert o["error"] == 0
    assert g["icm_moves"] > 32 * T
    assert_same(g, o, exact_energy=True)


def _streams(count, length, seed):
    x = O.Xoshiro(seed)
    out = []
    for _ in range(count):
        out.append(x.clone().words(length))
        x.jump()
    return out


@pytest.mark.parametrize("kind,kw,threads", [("2d_pm", {}, 2), ("rr3", dict(n=40), 2), ("rr3", dict(n=40), 1),
                                             ("qubo", dict(n=20, density=0.3), 2)])
def test_lockstep_bit_exact(gpu, kind, kw, threads):
    oi, gi, _ = make_pair(kind, seed=9, **kw)
    B, C, T = 6, 2, 12
    params = api.PtIcmParams(num_sweeps=T, warmup_fraction=0.25, beta=api.BetaOptions.Geometric(0.3, 3.0, B),
                             lo_num_beta=3, icm=True, num_replica_chains=C, threads=threads, sample=2,
                             sample_states=3, sample_limiting=0)
    runner = api.PtIcmRunner(gi, params)
    rng = np.random.default_rng(4)
    init = random_states(rng, C * B, oi.n).reshape(C, B, oi.n)
    ns = C * B if threads > 1 else 1
    per = T * (oi.n + 8) * (2 if threads > 1 else C * B + 2)
    streams = _streams(ns, per, 99)
    g = runner.run_lockstep(init, streams)
    o = O.pt_run(oi, runner.betas, T, num_chains=C, warmup_fraction=0.25, lo_beta_idx=runner.lo_beta_idx, icm=True,
                 threaded=threads > 1, sample=2, sample_states=3, sample_limiting=0, rng_mode=O.RNG_FEED,
                 streams=streams, init_states=init)
    assert o["error"] == 0
    assert_same(g, o, exact_energy=True)


@pytest.mark.parametrize("kind,kw,R,threads", [("qubo", dict(n=48, density=0.15), 100, 4), ("2d_gauss", {}, 64, 1),
                                               ("rr3", dict(n=70), 1, 1)])
def test_sa_bit_exact(gpu, kind, kw, R, threads):
    oi, gi, _ = make_pair(kind, seed=2, **kw)
    betas = np.linspace(0.1, 4.0, 40).astype(np.float32)
    params = api.SaParams(beta=api.BetaOptions.Arr(betas), num_replicas=R, threads=threads, seed=77)
    g = api.SaRunner(gi, params).run()
    o = O.sa_run(oi, betas, R, threads=threads, rng_mode=O.RNG_PHILOX, seed=77)
    assert g["R"] == o["R"]
    assert np.array_equal(g["final_states"], o["final_states"])
    assert np.array_equal(g["energies"], o["energies"])
    assert g["min_energy"] == o["min_energy"]


@pytest.mark.parametrize("kind,kw,R", [("rr3", dict(n=30000), 96), ("qubo", dict(n=21000, density=0.0004), 64)])
def test_sa_planes_in_global_memory(gpu, kind, kw, R):
    # instances whose 64 restarts' bit planes (8 bytes per site) do not fit in LDS: the planes stay
    # in the wave's global plane buffer, the row buffer in LDS
    oi, gi, _ = make_pair(kind, seed=4, **kw)
    betas = np.linspace(0.1, 4.0, 6).astype(np.float32)
    params = api.SaParams(beta=api.BetaOptions.Arr(betas), num_replicas=R, threads=1, seed=78)
    g = api.SaRunner(gi, params).run()
    o = O.sa_run(oi, betas, R, threads=1, rng_mode=O.RNG_PHILOX, seed=78)
    assert np.array_equal(g["final_states"], o["final_states"])
    assert np.array_equal(g["energies"], o["energies"])
    assert g["min_energy"] == o["min_energy"]


def test_batch_equals_single_and_plain(gpu):
    insts, params = [], []
    for k, n in enumerate([40, 100, 64, 200]):
        _, gi, _ = make_pair("rr3", seed=20 + k, n=n)
        insts.append(gi)
        params.append(api.PtIcmParams(num_sweeps=40, beta=api.BetaOptions.Geometric(0.2, 5.0, 8), lo_num_beta=4,
                                      sample=4, sample_states=8, seed=k))
    batch, st = api.run_pt_batch(insts, params)
    plain, _ = api.run_pt_batch(insts, params, plain=True)
    chunked, _ = api.run_pt_batch(insts, params, sweeps_per_launch=7, use_graph=True)
    tuned, tst = api.run_pt_batch(insts, params, autotune=True)
    assert tst["waves_per_group"] in (4, 8, 16) and tst["task_tiling"] in (1, 2)
    # contiguous task runs per wave, with more waves than some levels have tasks
    runs, rst = api.run_pt_batch(insts, params, waves_per_group=16, task_tiling=2, sweeps_per_launch=10,
                                 use_graph=True)
    assert rst["task_tiling"] == 2 and rst["launches"] == 4 and rst["graph"]
    strided, sst = api.run_pt_batch(insts, params, icm_growth=1)
    assert sst["icm_growth"] == 1
    teamed, tmst = api.run_pt_batch(insts, params, teams=3)
    assert tmst["teams"] == 3  # the three instances with at least 64 sites
    for k in range(len(insts)):
        single = api.PtIcmRunner(insts[k], params[k]).run()
        for other in (plain[k], chunked[k], tuned[k], runs[k], strided[k], teamed[k], single):
            for key in EXACT_KEYS + ["e", "gs_energies", "final_energies"]:
                assert np.array_equal(batch[k][key], other[key]), key
    assert st["attempts"] == sum(p.num_sweeps * i.n * 8 * 2 for p, i in zip(params, insts))


@pytest.mark.parametrize("kind,L,B,lo", [("2d_pm", 24, 48, 8), ("2d_gauss", 40, 20, 4), ("2d_pm", 64, 36, 4),
                                         ("2d_pm", 128, 12, 6), ("2d_pm", 25, 48, 8), ("2d_gauss", 57, 20, 4),
                                         ("2d_pm", 127, 12, 6), ("3d_gauss", 12, 48, 8), ("3d_pm", 16, 20, 4),
                                         ("3d_pm", 24, 12, 6), ("3d_gauss", 20, 12, 6)])
def test_lattice_bitboard_growth(gpu, kind, L, B, lo):
    # periodic L x L lattices (L >= 24; the last tiles along each axis partial unless 8 divides L)
    # and L x L x L lattices (L % 4 == 0, L >= 12) grow ICM clusters as 64-site bitboards (one
    # chunk of 32 pairs and a partial one at B - lo = 40; ICM arrays in LDS, or in HBM with the
    # spin planes parked at the largest sizes): the oracle's clusters, energies and samples
    oi, gi, _ = make_pair(kind, seed=120 + L, L=L)
    p = api.PtIcmParams(num_sweeps=6, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, B),
                        lo_num_beta=lo, icm=True, num_replica_chains=2, sample=2, sample_states=3,
                        sample_limiting=2, seed=1200 + L)
    runner = api.PtIcmRunner(gi, p)
    g = runner.run()
    o = O.pt_run(oi, runner.betas, 6, num_chains=2, warmup_fraction=0.5, lo_beta_idx=runner.lo_beta_idx, icm=True,
                 sample=2, sample_states=3, sample_limiting=2, rng_mode=O.RNG_PHILOX, seed=p.seed)
    assert o["error"] == 0
    assert g["icm_moves"] > 0
    assert_same(g, o, exact_energy=True, ctx=(kind, L))


@pytest.mark.parametrize("K", [3, 16])
def test_teams_with_icm_scratch_in_memory(gpu, K):
    # groups too large for their ICM scratch in LDS (spin planes in LDS, q and cluster words in
    # global memory, grown in the LDS the parked planes leave): each member keeps a private copy of
    # the scratch; same results as one work group and as the oracle, also over several launches
    cases = [("2d_pm", dict(L=100)), ("rr3", dict(n=12000))]
    insts, oracles, params = [], [], []
    for k, (kind, kw) in enumerate(cases):
        oi, gi, _ = make_pair(kind, seed=90 + k, **kw)
        insts.append(gi)
        oracles.append(oi)
        params.append(api.PtIcmParams(num_sweeps=8, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 12),
                                      lo_num_beta=6, icm=True, num_replica_chains=2, sample=2, sample_states=3,
                                      sample_limiting=2, seed=900 + k))
    alone, ast = api.run_pt_batch(insts, params, teams=1)
    assert ast["teams"] == 0
    team, tst = api.run_pt_batch(insts, params, teams=K)
    assert tst["teams"] == 2 and tst["work_groups"] >= 2 * K
    chunked, _ = api.run_pt_batch(insts, params, teams=K, sweeps_per_launch=3)
    for k, (oi, p) in enumerate(zip(oracles, params)):
        runner = api.PtIcmRunner(insts[k], p)
        o = O.pt_run(oi, runner.betas, 8, num_chains=2, warmup_fraction=0.5, lo_beta_idx=runner.lo_beta_idx,
                     icm=True, sample=2, sample_states=3, sample_limiting=2, rng_mode=O.RNG_PHILOX, seed=p.seed)
        assert o["error"] == 0
        assert_same(team[k], o, exact_energy=True, ctx=(K, k))
        for key in EXACT_KEYS + ["e", "gs_energies", "final_energies"]:
            assert np.array_equal(alone[k][key], team[k][key]), (K, k, key)
            assert np.array_equal(chunked[k][key], team[k][key]), (K, k, key)


@pytest.mark.parametrize("K", [2, 4, 16])
def test_teams_share_the_sweep(gpu, K):
    # a group's Metropolis sweep shared by K work groups that exchange spin words after every colour
    # class: the same results as one work group, also over several launches, against the oracle
    cases = [("3d_pm", dict(L=10)), ("rr3", dict(n=3000)), ("2d_pm", dict(L=24)), ("rr3_pm", dict(n=500))]
    insts, oracles, params = [], [], []
    for k, (kind, kw) in enumerate(cases):
        oi, gi, _ = make_pair(kind, seed=70 + k, **kw)
        insts.append(gi)
        oracles.append(oi)
        params.append(api.PtIcmParams(num_sweeps=12, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 12),
                                      lo_num_beta=6, icm=True, num_replica_chains=2 if k != 2 else 4,
                                      threads=1 if k != 3 else 2, sample=2, sample_states=3, sample_limiting=2,
                                      seed=700 + k))
    alone, ast = api.run_pt_batch(insts, params, teams=1)
    assert ast["teams"] == 0
    team, tst = api.run_pt_batch(insts, params, teams=K)
    assert tst["teams"] == 5 and tst["work_groups"] >= 2 * K
    chunked, _ = api.run_pt_batch(insts, params, teams=K, sweeps_per_launch=5)
    for k, (oi, p) in enumerate(zip(oracles, params)):
        runner = api.PtIcmRunner(insts[k], p)
        o = O.pt_run(oi, runner.betas, 12, num_chains=p.num_replica_chains, warmup_fraction=0.5,
                     lo_beta_idx=runner.lo_beta_idx, icm=True, threaded=p.threads > 1, sample=2, sample_states=3,
                     sample_limiting=2, rng_mode=O.RNG_PHILOX, seed=p.seed)
        assert o["error"] == 0
        assert_same(team[k], o, exact_energy=True, ctx=(K, k))
        for key in EXACT_KEYS + ["e", "gs_energies", "final_energies"]:
            assert np.array_equal(alone[k][key], team[k][key]), (K, k, key)
            assert np.array_equal(chunked[k][key], team[k][key]), (K, k, key)
