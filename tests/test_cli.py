"""The `tamc` command line (drop-in for src/main.rs + src/lib.rs:79-132) without a GPU: argument
and method-file parsing, beta resolution and error reporting through --dry-run."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tamc_amd import api, workloads as W

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "tamc")

README_PT = """---
PT:
  num_sweeps: 2000
  warmup_fraction: 0.5
  beta:
    Geometric:
      beta_min: 0.2
      beta_max: 5.0
      num_beta: 32
  lo_beta: 1.0
  icm: true
  num_replica_chains: 2
  threads: 1
  sample: 32
  sample_states: 32
  sample_limiting: 2
"""


@pytest.fixture(scope="module")
def files():
    d = tempfile.mkdtemp()
    i, j, K, n = W.lattice_2d(6, np.random.default_rng(3), pm_j=True)
    inst = os.path.join(d, "inst.txt")
    W.write_instance_file(inst, i, j, K)
    chi = os.path.join(d, "chi.txt")
    np.savetxt(chi, np.cos(np.arange(n)))
    return d, inst, chi


def run(args, method_text, d):
    m = os.path.join(d, "method.yml")
    with open(m, "w") as f:
        f.write(method_text)
    args = [a.replace("METHOD", m) for a in args]
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)


def test_readme_method_resolves_like_the_runner(files):
    d, inst, _ = files
    r = run(["--dry-run", "METHOD", inst, os.path.join(d, "o.yml")], README_PT, d)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    betas, lo = api.PtIcmParams(beta=api.BetaOptions.Geometric(0.2, 5.0, 32), lo_beta=1.0).resolve()
    assert np.array_equal(np.array(out["betas"], np.float32), betas) and out["lo_beta_idx"] == lo == 16
    assert out["n"] == 36 and out["num_sweeps"] == 2000 and out["sample_limiting"] == 2
    assert out["sample_output"] == "samples.bin" and out["sample_format"] == "bincode"


def test_flow_style_options_and_arguments(files):
    d, inst, chi = files
    text = "PT: {num_sweeps: 10, warmup_fraction: 0.25, beta: {Arr: [0.1, 0.5, 1.0, 2.0]}, lo_num_beta: 3, " \
           "icm: true, num_replica_chains: 4, threads: 8}\n"
    r = run(["--suscepts", chi, chi, "--sample-output", os.path.join(d, "s.pkl"), "METHOD", inst,
             os.path.join(d, "o.yml"), "--qubo", "--dry-run"], text, d)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["betas"] == [0.1, 0.5, 1.0, 2.0] and out["lo_beta_idx"] == 1
    assert out["sample"] is None and out["sample_states"] is None and out["sample_limiting"] is None
    assert out["nchi"] == 2 and out["qubo"] == 1 and out["threads"] == 8 and out["sample_format"] == "pickle"


def test_sa_method(files):
    d, inst, _ = files
    text = "SA:\n  beta:\n    Arr:\n      - 0.1\n      - 1.0\n      - 10.0\n  num_replicas: 64\n  threads: 4\n"
    r = run(["--dry-run", "METHOD", inst, os.path.join(d, "o.yml")], text, d)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["method"] == "SA" and out["betas"] == [0.1, 1.0, 10.0] and out["num_replicas"] == 64


@pytest.mark.parametrize("text,needle", [
    (README_PT.replace("  icm: true\n", ""), "missing field `icm`"),
    (README_PT.replace("beta_min: 0.2", "beta_min: 9.0"), "non-decreasing"),
    ("XX:\n  a: 1\n", "unknown variant `XX`"),
    ("GLA:\n  num_threads: 2\n", "GLA"),
    (README_PT.replace("num_sweeps: 2000", "num_sweeps: -5"), "num_sweeps"),
    (README_PT.replace("lo_beta: 1.0", "lo_beta: 1.0\n  lo_num_beta: 3").replace("icm: true", "icm: maybe"), "icm"),
])
def test_errors_are_reported_like_the_reference(files, text, needle):
    d, inst, _ = files
    r = run(["--dry-run", "METHOD", inst, os.path.join(d, "o.yml")], text, d)
    assert r.returncode == 1
    assert "tamc terminated with an error" in r.stderr and needle in r.stderr, r.stderr


def test_missing_instance_file(files):
    d, _, _ = files
    r = run(["--dry-run", "METHOD", os.path.join(d, "nope.txt"), os.path.join(d, "o.yml")], README_PT, d)
    assert r.returncode == 1 and "Unable to read adjancency" in r.stderr
