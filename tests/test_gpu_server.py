"""Serving loop (continuous batching): every request's result is identical to the same solve run
alone, whatever else shares the device and however its sweeps are sliced."""
import time

import numpy as np
import pytest

from tamc_amd import api
from tamc_amd._lib import TamcError
from tests.helpers import make_pair, random_states

pytestmark = pytest.mark.gpu

KEYS = ["e", "q", "suscept", "samples", "gs_steps", "gs_energies", "gs_states", "acceptance_counts", "final_states",
        "final_energies", "final_ids", "num_acceptances", "round_trips", "tags", "diffusion_hist"]


def same(a, b):
    for k in KEYS:
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k], b[k]), k
    assert a["icm_moves"] == b["icm_moves"]


def mixed_requests():
    specs = [("rr3", dict(n=40), 2, 1, 30), ("2d_pm", dict(L=8), 2, 1, 50), ("3d_gauss", dict(L=4), 4, 2, 40),
             ("rr3", dict(n=400), 2, 1, 60), ("qubo", dict(n=48, density=0.1), 3, 1, 35), ("rr3", dict(n=96), 1, 1, 45),
             ("2d_gauss", dict(L=12), 2, 2, 25), ("rr3_pm", dict(n=200), 2, 1, 70)]
    reqs = []
    for k, (kind, kw, C, threads, T) in enumerate(specs):
        _, gi, _ = make_pair(kind, seed=40 + k, **kw)
        p = api.PtIcmParams(num_sweeps=T, beta=api.BetaOptions.Geometric(0.2, 5.0, 8 + 4 * (k % 3)), lo_num_beta=4,
                            num_replica_chains=C, threads=threads, sample=4, sample_states=8, sample_limiting=k % 3,
                            seed=100 + k)
        reqs.append((gi, p))
    return reqs


def test_server_results_equal_single_runs(gpu):
    reqs = mixed_requests()
    alone = [api.PtIcmRunner(i, p).run() for i, p in reqs]
    init = random_states(np.random.default_rng(5), 2 * 8, reqs[0][0].n)
    alone_init = api.PtIcmRunner(*reqs[0]).run(init_states=init)
    # a short slice target forces many slices per request and different co-residents in each
    with api.Server(slice_ms=0.05) as srv:
        tickets = [srv.submit(i, p) for i, p in reqs[:4]]
        time.sleep(0.002)  # the rest arrive while the first ones run
        tickets += [srv.submit(i, p) for i, p in reqs[4:]]
        t_init = srv.submit(reqs[0][0], reqs[0][1], init_states=init)
        results = [srv.wait(t) for t in tickets]
        r_init, _ = srv.wait(t_init)
        st = srv.stats()
    for (res, times), ref in zip(results, alone):
        same(res, ref)
        assert times["total_ms"] >= times["queued_ms"] >= 0.0
        assert times["slices"] >= 1
    same(r_init, alone_init)
    assert st["completed"] == len(reqs) + 1 and st["failed"] == 0
    assert st["slices"] >= 2
    assert st["attempts"] == sum(r[0]["attempts"] for r in results) + r_init["attempts"]
    assert st["resident"] == 0 and st["pool_used"] == 0


def test_server_alone_and_in_the_mix(gpu):
    reqs = mixed_requests()
    target_i, target_p = reqs[3]
    with api.Server(slice_ms=0.1) as srv:
        alone, _ = srv.wait(srv.submit(target_i, target_p))
    with api.Server(slice_ms=0.1) as srv:
        others = [srv.submit(i, p) for i, p in reqs]
        t = srv.submit(target_i, target_p)
        mixed, _ = srv.wait(t)
        for o in others:
            srv.wait(o)
    same(alone, mixed)


def test_server_cancel_and_errors(gpu):
    reqs = mixed_requests()
    with api.Server(slice_ms=0.05, max_wave_slots=16) as srv:
        bad = api.PtIcmParams(num_sweeps=10, beta=api.BetaOptions.Arr([1.0, 0.5]), icm=False)
        with pytest.raises(TamcError):
            srv.submit(reqs[0][0], bad)
        with pytest.raises(TamcError):
            srv.poll(12345)
        tickets = [srv.submit(i, p) for i, p in reqs]
        srv.cancel(tickets[-1])  # still queued: only one request fits in 16 wave slots at a time
        with pytest.raises(TamcError):
            srv.poll(tickets[-1])
        done = [srv.wait(t) for t in tickets[:-1]]
        assert len(done) == len(reqs) - 1
        st = srv.stats()
        assert st["cancelled"] == 1 and st["completed"] == len(reqs) - 1
        assert st["pool_used"] == 0 and st["queued"] == 0


def _sa_requests():
    out = []
    for k, (kind, kw, R, threads, K) in enumerate([("qubo", dict(n=200, density=0.1), 300, 3, 60),
                                                   ("2d_gauss", dict(L=6), 64, 1, 30),
                                                   ("qubo", dict(n=2048, density=0.05), 1024, 1, 24),
                                                   ("rr3", dict(n=70), 1, 1, 40)]):
        _, gi, _ = make_pair(kind, seed=60 + k, **kw)
        betas = np.linspace(0.1, 4.0, K).astype(np.float32)
        out.append((gi, api.SaParams(beta=api.BetaOptions.Arr(betas), num_replicas=R, threads=threads, seed=70 + k)))
    return out


def test_server_sa_requests_equal_single_runs(gpu):
    # SA solves served beside PT-ICM solves: a short slice target splits each SA schedule over
    # several launches (the restarts' planes stay on the device in between); every result equals
    # tamc_sa_run on the same inputs, and the PT results stay those of their single runs
    sa = _sa_requests()
    pt = mixed_requests()[:4]
    sa_alone = [api.SaRunner(i, p).run() for i, p in sa]
    pt_alone = [api.PtIcmRunner(i, p).run() for i, p in pt]
    with api.Server(slice_ms=0.5) as srv:
        ts = [srv.submit_sa(i, p) for i, p in sa[:2]]
        tp = [srv.submit(i, p) for i, p in pt]
        ts += [srv.submit_sa(i, p) for i, p in sa[2:]]
        with pytest.raises(TamcError):
            srv.poll(ts[0])  # an SA ticket is polled with poll_sa
        sa_served = [srv.wait_sa(t) for t in ts]
        pt_served = [srv.wait(t)[0] for t in tp]
        st = srv.stats()
    for (res, times), ref in zip(sa_served, sa_alone):
        assert res["R"] == ref["R"]
        assert np.array_equal(res["final_states"], ref["final_states"])
        assert np.array_equal(res["energies"], ref["energies"])
        assert res["min_energy"] == ref["min_energy"] and res["attempts"] == ref["attempts"]
    assert sa_served[2][1]["slices"] > 1  # the 2048-site QUBO took several launches
    for res, ref in zip(pt_served, pt_alone):
        same(res, ref)
    assert st["completed"] == len(sa) + len(pt) and st["failed"] == 0 and st["pool_used"] == 0
