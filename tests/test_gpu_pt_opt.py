"""tamc-pt-opt on the GPU (src/bin/tamc-pt-opt.rs): the command line and the Python optimiser make
the same library calls, so on one seed they produce the same temperature sets and round-trip times;
the optimised set keeps its end points, increases, and the sweeps double."""
import os
import subprocess
import tempfile

import numpy as np
import pytest
import yaml

from tamc_amd import api, workloads as W

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "tamc-pt-opt")

PT_TEXT = """PT:
  num_sweeps: 400
  warmup_fraction: 0.5
  beta:
    Geometric:
      beta_min: 0.2
      beta_max: 5.0
      num_beta: 12
  lo_beta: 1.0
  icm: true
  num_replica_chains: 2
  threads: 1
"""


def test_pt_opt_cli_matches_python(gpu):
    d = tempfile.mkdtemp()
    files = []
    for k in range(2):
        i, j, K, n = W.lattice_2d(6, np.random.default_rng(40 + k), pm_j=True)
        files.append(os.path.join(d, "inst%d.txt" % k))
        W.write_instance_file(files[-1], i, j, K)
    with open(os.path.join(d, "pt.yml"), "w") as f:
        f.write(PT_TEXT)
    seed = 5150
    r = subprocess.run([CLI, "--max-iters", "3", "--tolerance", "0.0", "pt.yml"] + files, capture_output=True,
                       text=True, cwd=d, timeout=300, env=dict(os.environ, TAMC_SEED=str(seed)))
    assert r.returncode == 0, r.stderr
    assert "Final temperature array" in r.stdout and "Writing to opt_params.yml" in r.stdout
    with open(os.path.join(d, "opt_params.yml")) as f:
        out = yaml.safe_load(f)["PT"]
    got = np.array(out["beta"]["Arr"], np.float32)
    assert out["num_sweeps"] == 800 and len(got) == 12
    assert got[0] == np.float32(0.2) and abs(got[-1] - 5.0) < 1e-5 and np.all(np.diff(got) > 0)
    with open(os.path.join(d, "tau_hist.csv")) as f:
        rows = [[np.float32(x) for x in line.strip().rstrip(",").split(",")] for line in f if line.strip()]
    assert len(rows) == 3 and all(len(row) == 2 for row in rows)

    params = api.PtIcmParams(num_sweeps=400, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 12),
                             lo_beta=1.0, icm=True, num_replica_chains=2, threads=1, seed=seed)
    insts = [api.BqmIsingInstance.from_instance_file(f) for f in files]
    opt, taus = api.pt_optimize_beta(insts, params, 3, 0.2, 0.85, 0.0)
    assert np.array_equal(np.array(opt.beta.arr, np.float32), got)
    assert np.array_equal(taus, np.array(rows, np.float32))
    assert opt.num_sweeps == 800
