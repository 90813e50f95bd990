"""GPU checks of the boundary additions: an instance created from caller CSR arrays runs exactly
like the file-built one, and a run resumed from a full PTState (tamc_pt_run_resume, the
counterpart of PtIcmRunner::run(Some(states)), src/pt.rs:397-408) matches the oracle resumed from
the same state."""
import numpy as np
import pytest

import oracle as O
from tamc_amd import api
from tests.helpers import make_pair
from tests.test_gpu_parity import EXACT_KEYS, assert_same

pytestmark = pytest.mark.gpu


def _params(B=8, T=40, C=2, threads=1, seed=3, lo=4):
    return api.PtIcmParams(num_sweeps=T, warmup_fraction=0.25, beta=api.BetaOptions.Geometric(0.2, 5.0, B),
                           lo_num_beta=B - lo, icm=True, num_replica_chains=C, threads=threads, sample=4,
                           sample_states=4, sample_limiting=0, seed=seed)


@pytest.mark.parametrize("kind,kw", [("2d_pm", {}), ("qubo", dict(n=30, density=0.2)), ("rr3", dict(n=80))])
def test_created_instance_runs_like_file_build(gpu, kind, kw):
    oi, gi, _ = make_pair(kind, seed=5, **kw)
    row_ptr, col, J, bias = gi.csr()
    w = np.random.default_rng(1).standard_normal((2, gi.n))
    gc = api.BqmIsingInstance.from_csr(row_ptr, col, J, bias=bias, offset=gi.offset, suscept=w, qubo=gi.qubo)
    gi.with_suscept(w)
    p = _params()
    a = api.PtIcmRunner(gi, p).run()
    b = api.PtIcmRunner(gc, p).run()
    for key in EXACT_KEYS + ["e", "gs_energies", "final_energies", "suscept"]:
        assert np.array_equal(a[key], b[key]), key


@pytest.mark.parametrize("kind,kw,C,threads", [("2d_pm", {}, 2, 1), ("rr3_pm", dict(n=64), 4, 2),
                                               ("qubo", dict(n=24, density=0.2), 2, 1)])
def test_resume_from_full_state(gpu, kind, kw, C, threads):
    oi, gi, _ = make_pair(kind, seed=8, **kw)
    p1 = _params(C=C, threads=threads, T=30, seed=11)
    r1 = api.PtIcmRunner(gi, p1)
    first = r1.run()
    assert first["round_trips"].sum() + first["num_acceptances"].sum() > 0
    p2 = _params(C=C, threads=threads, T=50, seed=12)
    r2 = api.PtIcmRunner(gi, p2)
    g = r2.run_from(first)
    state = dict(energies=first["final_energies"], num_acceptances=first["num_acceptances"],
                 round_trips=first["round_trips"], tags=first["tags"], diffusion_hist=first["diffusion_hist"])
    o = O.pt_run(oi, r2.betas, 50, num_chains=C, warmup_fraction=0.25, lo_beta_idx=r2.lo_beta_idx, icm=True,
                 threaded=threads > 1, sample=4, sample_states=4, sample_limiting=0, rng_mode=O.RNG_PHILOX,
                 seed=12, init_states=first["final_states"], init_state=state)
    assert o["error"] == 0
    assert_same(g, o, exact_energy=True)
    # the carried-over counters are part of the totals (count_acc, src/pt.rs:585-591)
    assert np.all(g["num_acceptances"] >= first["num_acceptances"])
    assert np.array_equal(g["acceptance_counts"], g["num_acceptances"].sum(axis=0))
    assert np.all(g["round_trips"] >= first["round_trips"])


def test_resume_with_states_only_equals_init_states(gpu):
    oi, gi, _ = make_pair("2d_pm", seed=2)
    rng = np.random.default_rng(3)
    init = (rng.integers(0, 2, size=(2, 8, gi.n)) * 2 - 1).astype(np.int8)
    r = api.PtIcmRunner(gi, _params(T=20))
    a = r.run(init_states=init)
    b = r.run_from({"states": init})
    for key in EXACT_KEYS + ["e", "gs_energies", "final_energies"]:
        assert np.array_equal(a[key], b[key]), key
