"""GLA on the GPU (src/gla.rs:122-196): the partitions' sub-instances run in one batched launch.
Each partition energy and ground state equals a single PtIcmRunner run of that sub-instance with
seed + k; boundary energies equal the restatement in oracle/gla.py on the joined state; and the total
equals the full instance's energy_ref of the joined state (exact with +-J couplings)."""
import dataclasses
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from oracle import gla as OG
from tamc_amd import api, samples, workloads as W

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "tamc")


def _params(seed):
    return api.PtIcmParams(num_sweeps=300, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 16),
                           lo_beta=1.0, icm=True, num_replica_chains=2, threads=1, sample=8, sample_states=16,
                           sample_limiting=1, seed=seed)


def _unpack_u64(words, n):
    bits = np.unpackbits(np.asarray(words, np.uint64).view(np.uint8), bitorder="little")[:n]
    return (1 - 2 * bits.astype(np.int8)).astype(np.int8)


@pytest.mark.parametrize("kind", ["2d_pm", "rr3_pm"])
def test_gla_matches_sub_solves_and_restatement(gpu, kind):
    rng = np.random.default_rng(21)
    if kind == "2d_pm":
        i, j, K, n = W.lattice_2d(10, rng, pm_j=True)
    else:
        i, j, K, n = W.random_regular3(120, rng, gaussian=False, bias_sigma=0.0)
    inst = api.BqmIsingInstance.from_lines(i, j, K)
    oi = O.Instance.from_lines(i, j, K)
    perm = rng.permutation(n)
    cuts = [0, n // 4, n // 2, (3 * n) // 4, n]
    parts = [np.sort(perm[a:b]).astype(np.uint32) for a, b in zip(cuts[:-1], cuts[1:])]
    params = _params(1234)
    g = api.run_gla(inst, params, partitions=parts)
    assert g["num_partitions"] == 4 and g["instance_size"] == n
    for k, p in enumerate(parts):
        sub = inst.induced_subgraph(p)
        ref = api.PtIcmRunner(sub, dataclasses.replace(params, seed=params.seed + k)).run()
        assert g["partition_energies"][k] == ref["gs_energies"][-1]
        assert np.array_equal(g["joined_state"][p], _unpack_u64(ref["gs_states"][-1], len(p)))
    bnd, total_bnd = OG.boundary_energies(oi.row_ptr, oi.col, oi.J, [list(map(int, p)) for p in parts],
                                          g["joined_state"])
    assert [(a, b) for a, b, _ in g["boundary_energies"]] == [(a, b) for a, b, _ in bnd]
    assert np.array_equal(g["boundary_energies_f32"], np.array([e for _, _, e in bnd], np.float32))
    assert g["boundary_energy"] == total_bnd
    assert g["total_energy"] == np.float32(g["bulk_energy"] + g["boundary_energy"])
    assert g["total_energy"] == np.float32(oi.energy(g["joined_state"]))


def test_cli_gla_output_matches_api(gpu):
    d = tempfile.mkdtemp()
    i, j, K, n = W.lattice_2d(8, np.random.default_rng(3), pm_j=True)
    W.write_instance_file(os.path.join(d, "inst.txt"), i, j, K)
    with open(os.path.join(d, "parts.txt"), "w") as f:
        f.write("\n".join(" ".join(str(r * 8 + c) for r in range(4 * h, 4 * h + 4) for c in range(8))
                          for h in range(2)) + "\n")
    with open(os.path.join(d, "gla.yml"), "w") as f:
        f.write("GLA:\n  num_threads: 2\n  pt_params:\n    num_sweeps: 300\n    warmup_fraction: 0.5\n"
                "    beta:\n      Geometric:\n        beta_min: 0.2\n        beta_max: 5.0\n        num_beta: 16\n"
                "    lo_beta: 1.0\n    icm: true\n    num_replica_chains: 2\n    threads: 1\n    sample: 8\n"
                "    sample_states: 16\n    sample_limiting: 1\n  partition_file: parts.txt\n")
    seed = 99
    for out in ("out.bin", "out.pkl"):
        r = subprocess.run([CLI, "gla.yml", "inst.txt", out], capture_output=True, text=True, cwd=d, timeout=120,
                           env=dict(os.environ, TAMC_SEED=str(seed)))
        assert r.returncode == 0, r.stderr
        assert "Total energy:" in r.stdout and "tamc finished successfully" in r.stdout
        got = samples.read_gla_results(os.path.join(d, out))
        inst = api.BqmIsingInstance.from_instance_file(os.path.join(d, "inst.txt"))
        want = api.run_gla(inst, _params(seed), partition_file=os.path.join(d, "parts.txt"))
        assert got["instance_size"] == 64 and got["num_partitions"] == 2
        assert np.array_equal(got["partition_energies"], want["partition_energies"])
        assert [(a, b) for a, b, _ in got["boundary_energies"]] == [(a, b) for a, b, _ in want["boundary_energies"]]
        assert np.array_equal(np.array([e for _, _, e in got["boundary_energies"]], np.float32),
                              want["boundary_energies_f32"])
        assert got["total_energy"] == want["total_energy"] and got["final_state"].size == 0
