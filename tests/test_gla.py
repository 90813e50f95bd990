"""GLA pieces that run without a GPU: the partition file reader (src/gla.rs:35-59,
src/util.rs:142-172), the induced sub-instance (src/ising.rs:327-355), the boundary-energy
restatement (src/gla.rs:73-120,175-190), the GlaResults writer (bincode and pickle) and the GLA
method file through the command line's --dry-run."""
import ctypes as C
import json
import os
import pickle
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from oracle import gla as OG
from tamc_amd import _lib, api, workloads as W
from tests.helpers import random_states

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "tamc")

PART_TEXT = "0 1 2 3 4 5 6 7\n\n  8 9 x 10\nfoo 11\n\t15 14 13 12 11 10x\n"


def _write(d, name, text):
    p = os.path.join(d, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def test_partition_reader_matches_restatement():
    i, j, K, n = W.lattice_2d(4, np.random.default_rng(1), pm_j=True)
    inst = api.BqmIsingInstance.from_lines(i, j, K)
    d = tempfile.mkdtemp()
    got = api.read_partition(inst, _write(d, "p.txt", PART_TEXT))
    want = OG.read_graph_partition(PART_TEXT, n)
    # hand-worked: "8 9 x 10" stops at x, "foo 11" is skipped, "10x" reads 10
    assert want == [[0, 1, 2, 3, 4, 5, 6, 7], [8, 9], [10, 11, 12, 13, 14, 15]]
    assert [list(map(int, p)) for p in got] == want


@pytest.mark.parametrize("text,msg", [
    ("0 1 2\n3 4 2\n", "Index 2 specified more than once in partition 1"),
    ("0 1 16\n", "Partitition must be specified for instance size 16. Found index 16"),
])
def test_partition_reader_errors(text, msg):
    i, j, K, n = W.lattice_2d(4, np.random.default_rng(1), pm_j=True)
    inst = api.BqmIsingInstance.from_lines(i, j, K)
    d = tempfile.mkdtemp()
    with pytest.raises(_lib.TamcError, match=msg):
        api.read_partition(inst, _write(d, "p.txt", text))
    with pytest.raises(ValueError, match=msg):
        OG.read_graph_partition(text, n)


@pytest.mark.parametrize("qubo", [False, True])
def test_induced_subgraph_matches_restatement(qubo):
    rng = np.random.default_rng(5)
    if qubo:
        i, j, K, n = W.qubo_sparse(30, 0.25, rng)
    else:
        i, j, K, n = W.lattice_2d(6, rng, pm_j=False)
    inst = api.BqmIsingInstance.from_lines(i, j, K, qubo=qubo)
    oi = O.Instance.from_lines(i, j, K, qubo=qubo)
    nodes = np.sort(rng.choice(n, size=n // 2, replace=False)).astype(np.uint32)
    sub = inst.induced_subgraph(nodes)
    rp, col, J, bias = sub.csr()
    wrp, wcol, wJ, wbias = OG.induced_subgraph(oi.row_ptr, oi.col, oi.J, oi.bias, nodes)
    assert sub.n == len(nodes) and sub.offset == 0.0
    assert np.array_equal(rp, wrp) and np.array_equal(col, wcol)
    assert np.array_equal(J, wJ) and np.array_equal(bias, wbias)
    # the ICM graph is the pattern of the restricted rows (QUBO self entries included)
    grp, gcol = sub.graph()
    last = int(np.max(np.concatenate([np.repeat(np.arange(len(nodes)), np.diff(wrp)), wcol]))) + 1 if len(wcol) else 0
    assert len(grp) == last + 1
    assert np.array_equal(grp, wrp[:last + 1]) and np.array_equal(gcol, wcol)
    if qubo:
        assert any(c == r for r in range(len(nodes)) for c in wcol[wrp[r]:wrp[r + 1]])


def test_boundary_restatement_pins_energy_split():
    """energy_ref(full) == sum over partitions of energy_ref(sub) + sum of boundary energies, for
    any state (exact with +-J couplings and integer biases)."""
    rng = np.random.default_rng(7)
    i, j, K, n = W.lattice_2d(8, rng, pm_j=True)
    i = np.concatenate([i, np.arange(n)])  # integer biases on every site
    j = np.concatenate([j, np.arange(n)])
    K = np.concatenate([K, rng.integers(-2, 3, n).astype(np.float64)])
    oi = O.Instance.from_lines(i, j, K)
    perm = rng.permutation(n)
    parts = [sorted(int(x) for x in perm[a:b]) for a, b in ((0, 20), (20, 41), (41, 64))]
    for s in random_states(rng, 5, n):
        _, total_bnd = OG.boundary_energies(oi.row_ptr, oi.col, oi.J, parts, s)
        bulk = np.float32(0.0)
        for p in parts:
            rp, col, J, bias = OG.induced_subgraph(oi.row_ptr, oi.col, oi.J, oi.bias, p)
            sub = O.Instance.from_lines(
                np.concatenate([np.repeat(np.arange(len(p)), np.diff(rp)), np.arange(len(p))]),
                np.concatenate([col, np.arange(len(p))]), np.concatenate([J, bias]).astype(np.float64))
            bulk = np.float32(bulk + np.float32(sub.energy(s[p])))
        assert np.float32(bulk + total_bnd) == np.float32(oi.energy(s))
    # every coupling between partitions is counted once
    es = OG.partition_graph(oi.row_ptr, oi.col, oi.J, parts)
    assert sum(len(e) for _, e in es) == sum(1 for r in range(n) for c in oi.col[oi.row_ptr[r]:oi.row_ptr[r + 1]]
                                            if c > r and any(r in p and c not in p for p in parts))


def _gla_result(np_, pe, pairs, be, total, n):
    R = _lib.GlaResult()
    keep = dict(pe=np.array(pe, np.float32), bp=np.array(pairs, np.uint32).reshape(-1),
                be=np.array(be, np.float32), js=np.ones(n, np.int8))
    R.instance_size, R.num_partitions, R.num_boundaries = n, np_, len(be)
    R.partition_energies = keep["pe"].ctypes.data_as(_lib._f32p)
    R.boundary_pairs = keep["bp"].ctypes.data_as(_lib._u32p)
    R.boundary_energies = keep["be"].ctypes.data_as(_lib._f32p)
    R.total_energy = total
    R.joined_state = keep["js"].ctypes.data_as(_lib._i8p)
    return R, keep


def test_gla_results_writer_layouts():
    R, keep = _gla_result(3, [-10.5, -7.0, -3.25], [(0, 1), (1, 2), (2, 0)], [1.5, -2.0, 0.25], -21.0, 64)
    d = tempfile.mkdtemp()
    b = os.path.join(d, "out.bin")
    _lib.check(_lib.lib().tamc_write_gla_results(C.byref(R), b.encode()))
    want = struct.pack("<IIQ3fQ", 64, 3, 3, -10.5, -7.0, -3.25, 3)
    for (a, c), e in zip([(0, 1), (1, 2), (2, 0)], [1.5, -2.0, 0.25]):
        want += struct.pack("<IIf", a, c, e)
    want += struct.pack("<Qf", 0, -21.0)
    assert open(b, "rb").read() == want
    from tamc_amd import samples
    for path in (b, b[:-4] + ".pkl"):
        if path.endswith(".pkl"):
            _lib.check(_lib.lib().tamc_write_gla_results(C.byref(R), path.encode()))
        g = samples.read_gla_results(path)
        assert g["instance_size"] == 64 and g["num_partitions"] == 3 and g["total_energy"] == np.float32(-21.0)
        assert list(g["partition_energies"]) == [-10.5, -7.0, -3.25] and g["final_state"].size == 0
        assert g["boundary_energies"] == [(0, 1, 1.5), (1, 2, -2.0), (2, 0, 0.25)]
    p = os.path.join(d, "out.pkl")
    _lib.check(_lib.lib().tamc_write_gla_results(C.byref(R), p.encode()))
    with open(p, "rb") as f:
        obj = pickle.load(f)  # written by this build's library
    assert obj == {"instance_size": 64, "num_partitions": 3, "partition_energies": [-10.5, -7.0, -3.25],
                   "boundary_energies": [(0, 1, 1.5), (1, 2, -2.0), (2, 0, 0.25)], "final_state": [],
                   "total_energy": -21.0}
    assert list(obj) == ["instance_size", "num_partitions", "partition_energies", "boundary_energies",
                         "final_state", "total_energy"]


GLA_TEXT = """GLA:
  num_threads: 2
  pt_params:
    num_sweeps: 200
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
  partition_file: parts.txt
"""


def test_cli_gla_dry_run():
    d = tempfile.mkdtemp()
    i, j, K, n = W.lattice_2d(8, np.random.default_rng(3), pm_j=True)
    W.write_instance_file(os.path.join(d, "inst.txt"), i, j, K)
    _write(d, "parts.txt", "\n".join(" ".join(str(r * 8 + c) for r in range(4 * h, 4 * h + 4) for c in range(8))
                                     for h in range(2)) + "\n")
    _write(d, "gla.yml", GLA_TEXT)
    r = subprocess.run([CLI, "--dry-run", "gla.yml", "inst.txt", "out.pkl"], capture_output=True, text=True, cwd=d,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    cfg = json.loads(r.stdout)
    assert cfg["method"] == "GLA" and cfg["n"] == 64 and cfg["num_partitions"] == 2
    assert cfg["lo_beta_idx"] == 8 and cfg["output_format"] == "pickle"
    _write(d, "parts.txt", "0 1 2\n2\n")
    r = subprocess.run([CLI, "--dry-run", "gla.yml", "inst.txt", "out.pkl"], capture_output=True, text=True, cwd=d,
                       timeout=60)
    assert r.returncode == 1 and "Index 2 specified more than once in partition 1" in r.stderr
