"""The `tamc` command line end to end on the GPU: the files it writes hold the same results as the
runner API on the same seed (TAMC_SEED), in the reference's formats."""
import os
import subprocess
import tempfile

import numpy as np
import pytest
import yaml

from tamc_amd import api, samples, workloads as W

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "tamc")

PT_TEXT = """---
PT:
  num_sweeps: 300
  warmup_fraction: 0.5
  beta:
    Geometric:
      beta_min: 0.2
      beta_max: 5.0
      num_beta: 16
  lo_beta: 1.0
  icm: true
  num_replica_chains: 2
  threads: 1
  sample: 8
  sample_states: 16
  sample_limiting: 1
"""


def cli(args, d, seed):
    env = dict(os.environ, TAMC_SEED=str(seed))
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120, cwd=d, env=env)


def test_cli_pt_outputs_match_runner(gpu):
    d = tempfile.mkdtemp()
    i, j, K, n = W.lattice_2d(8, np.random.default_rng(31), pm_j=True)
    inst_path = os.path.join(d, "inst.txt")
    W.write_instance_file(inst_path, i, j, K)
    with open(os.path.join(d, "pt.yml"), "w") as f:
        f.write(PT_TEXT)
    seed = 4242
    r = cli(["pt.yml", "inst.txt", "out.yml"], d, seed)
    assert r.returncode == 0, r.stderr
    assert "tamc finished successfully" in r.stdout and "** Ground state energy **" in r.stdout
    r2 = cli(["--sample-output", "s.pkl", "pt.yml", "inst.txt", "out2.yml"], d, seed)
    assert r2.returncode == 0, r2.stderr

    params = api.PtIcmParams(num_sweeps=300, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 16),
                             lo_beta=1.0, icm=True, num_replica_chains=2, threads=1, sample=8, sample_states=16,
                             sample_limiting=1, seed=seed)
    inst = api.BqmIsingInstance.from_instance_file(inst_path)
    ref = api.PtIcmRunner(inst, params).run()

    with open(os.path.join(d, "out.yml")) as f:
        y = yaml.safe_load(f)
    assert y["params"]["beta"] == {"Geometric": {"beta_min": 0.2, "beta_max": 5.0, "num_beta": 16}}
    assert y["params"]["lo_num_beta"] is None and y["params"]["sample_limiting"] == 1
    assert y["gs_time_steps"] == [int(x) for x in ref["gs_steps"]]
    assert np.array_equal(np.array(y["gs_energies"], np.float32), ref["gs_energies"])
    assert y["gs_states"] == [[int(w) for w in row] for row in ref["gs_states"]]
    assert y["acceptance_counts"] == [int(x) for x in ref["acceptance_counts"]]
    assert y["instance_size"] == 64 and y["num_measurements"] == 0 and y["timing"] > 0

    b = samples.read_thermal_samples(os.path.join(d, "samples.bin"))
    p = samples.read_thermal_samples(os.path.join(d, "s.pkl"))  # written by this build's CLI
    for obj in (b, p):
        assert obj["instance_size"] == 64
        assert np.array_equal(obj["beta_arr"], params.resolve()[0])
        assert np.array_equal(np.array(obj["e"], np.float32), ref["e"])
        assert np.array_equal(np.array(obj["q"], np.int32), ref["q"])
        assert np.array_equal(np.array(obj["samples"], np.uint8), ref["samples"])


def test_cli_sa_outputs_match_runner(gpu):
    d = tempfile.mkdtemp()
    i, j, K, n = W.qubo_sparse(40, 0.2, np.random.default_rng(32))
    inst_path = os.path.join(d, "q.txt")
    W.write_instance_file(inst_path, i, j, K)
    betas = [round(0.1 + 0.05 * k, 2) for k in range(60)]
    with open(os.path.join(d, "sa.yml"), "w") as f:
        f.write("SA:\n  beta:\n    Arr: [%s]\n  num_replicas: 96\n  threads: 8\n" % ", ".join(map(str, betas)))
    seed = 77
    r = cli(["--qubo", "sa.yml", "q.txt", "out.yml"], d, seed)
    assert r.returncode == 0, r.stderr
    inst = api.BqmIsingInstance.from_instance_file(inst_path, qubo=True)
    ref = api.SaRunner(inst, api.SaParams(beta=api.BetaOptions.Arr(betas), num_replicas=96, threads=8,
                                          seed=seed)).run()
    with open(os.path.join(d, "out.yml")) as f:
        y = yaml.safe_load(f)
    assert np.float32(y["min_energy"]) == ref["min_energy"]
    assert np.array_equal(np.array(y["energies"], np.float32), ref["energies"])
    assert y["params"]["num_replicas"] == 96 and y["params"]["threads"] == 8
    a = samples.read_anneal_state(os.path.join(d, "samples.bin"))
    ngs = int(np.sum(ref["energies"] <= ref["min_energy"] + np.float32(1e-8)))
    assert a["instance_size"] == 40 and a["num_replicas"] == ngs and a["bytes"].size == ngs * 5
    # the packed ground states decode to states whose energy_ref is the minimum energy
    states = samples.unpack_states(a["bytes"], 40)
    assert np.array_equal(inst.energy(states), np.full(ngs, ref["min_energy"], np.float32))
