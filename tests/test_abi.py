"""CPU checks of the C ABI boundary: the library loads, exports every function include/tamc_gpu.h
declares, and its host-side parts (instance parsing/building, PtIcmRunner::new resolution, bincode
writers) agree with the CPU restatement.  No kernel is launched here."""
import ctypes as C
import os
import re
import struct
import tempfile

import numpy as np
import pytest

import oracle as O
from tamc_amd import _lib, api, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tamc_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tamc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) <= set(names)


def _write(lines):
    f = tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False)
    f.write("\n".join(lines) + "\n")
    f.close()
    return f.name


@pytest.mark.parametrize("qubo", [False, True])
def test_instance_file_parity_with_restatement(qubo):
    rng = np.random.default_rng(4)
    i, j, K, n = W.qubo_sparse(20, 0.3, rng) if qubo else W.random_regular3(40, rng, bias_sigma=0.2)
    lines = ["%d %d %s" % (a, b, np.format_float_positional(np.float32(k), trim="0")) for a, b, k in zip(i, j, K)]
    lines += ["# comment line", "3 4", "1 2 1e-3", "7 7 -0.5junk"]
    path = _write(lines)
    try:
        g = api.BqmIsingInstance.from_instance_file(path, qubo=qubo)
        o = O.Instance.from_file(path, qubo=qubo)
    finally:
        os.unlink(path)
    row_ptr, col, J, bias = g.csr()
    assert g.n == o.n and np.array_equal(row_ptr, o.row_ptr) and np.array_equal(col, o.col)
    assert np.array_equal(J, o.J) and np.array_equal(bias, o.bias) and g.offset == o.offset
    assert g.graph_nodes == o.g_nodes


def test_missing_instance_file_is_io_error():
    with pytest.raises(_lib.TamcError) as e:
        api.BqmIsingInstance.from_instance_file("/nonexistent/instance.txt")
    assert e.value.status == _lib.TAMC_E_IO


def test_resolve_matches_restatement_and_reference_rules():
    p = api.PtIcmParams(beta=api.BetaOptions.Geometric(0.2, 5.0, 32), lo_beta=1.0)
    b, lo = p.resolve()
    assert np.array_equal(b, O.beta_geometric(0.2, 5.0, 32)) and lo == 16
    p = api.PtIcmParams(beta=api.BetaOptions.Geometric(0.1, 5.0, 48), lo_beta=1.0)
    assert p.resolve()[1] == int(np.argmax(O.beta_geometric(0.1, 5.0, 48) >= 1.0))
    # lo_num_beta takes priority and is clamped to the ladder
    assert api.PtIcmParams(beta=api.BetaOptions.Geometric(0.2, 5.0, 8), lo_beta=1.0, lo_num_beta=3).resolve()[1] == 5
    assert api.PtIcmParams(beta=api.BetaOptions.Geometric(0.2, 5.0, 8), lo_num_beta=100).resolve()[1] == 0
    # lo_beta beyond the ladder -> largest beta
    assert api.PtIcmParams(beta=api.BetaOptions.Geometric(0.2, 5.0, 8), lo_beta=9.0).resolve()[1] == 7
    # icm without lo_beta/lo_num_beta and a decreasing ladder are errors (reference panics)
    with pytest.raises(_lib.TamcError):
        api.PtIcmParams(beta=api.BetaOptions.Geometric(0.2, 5.0, 8)).resolve()
    with pytest.raises(_lib.TamcError):
        api.PtIcmParams(beta=api.BetaOptions.Arr([1.0, 0.5]), icm=False).resolve()
    b, lo = api.PtIcmParams(beta=api.BetaOptions.Arr([0.5, 0.5, 2.0]), icm=False).resolve()
    assert list(b) == [0.5, 0.5, 2.0] and lo == 0


def _fake_result(B=3, C=2, n=10, nm=2, nchi=1, slots=1, sev=2):
    np_ = C // 2
    keep = {}
    R = _lib.PtResult()
    R.n, R.B, R.C, R.npairs, R.n_meas, R.nchi = n, B, C, np_, nm, nchi
    R.n_samp_slots, R.n_samp_events, R.nbytes = slots, sev, (n + 7) // 8

    def put(name, arr, ctype):
        a = np.ascontiguousarray(arr)
        keep[name] = a
        setattr(R, name, a.ctypes.data_as(ctype))
        return a

    e = put("e", np.arange(B * nm * C, dtype=np.float32) * 0.5, _lib._f32p)
    q = put("q", np.arange(B * nm * np_, dtype=np.int32) - 3, _lib._i32p)
    chi = put("suscept", np.arange(B * nchi * nm * np_, dtype=np.float32) + 0.25, _lib._f32p)
    smp = put("samples", np.arange(slots * sev * C * R.nbytes, dtype=np.uint8), _lib._u8p)
    return R, keep, (e, q, chi, smp)


def test_thermal_samples_bincode_layout():
    R, keep, (e, q, chi, smp) = _fake_result()
    betas = np.array([0.5, 1.0, 2.0], np.float32)
    path = tempfile.mktemp(suffix=".bin")
    _lib.check(_lib.lib().tamc_write_thermal_samples(C.byref(R), betas.ctypes.data_as(_lib._f32p), 2, path.encode()))
    data = open(path, "rb").read()
    os.unlink(path)
    # independent decoder of the bincode-1 layout of PtIcmThermalSamples (src/pt.rs:116-125)
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, data, off)
        off += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    assert take("B") == 2 and take("Q") == 10
    nb = take("Q")
    assert list(take("%df" % nb)) == list(betas)
    ns = take("Q")
    samples = []
    for _ in range(ns):
        m = take("Q")
        samples.append([bytes(take("%dB" % take("Q"))) for _ in range(m)])
    assert len(samples) == 1 and len(samples[0]) == 4 and samples[0][1] == bytes(smp[2:4])
    assert take("Q") == 3
    ee = [list(take("%df" % take("Q"))) for _ in range(3)]
    assert ee == [list(e[k * 4:(k + 1) * 4]) for k in range(3)]
    assert take("Q") == 3
    qq = [take("%di" % take("Q")) for _ in range(3)]
    assert [list(np.atleast_1d(x)) for x in qq] == [list(q[k * 2:(k + 1) * 2]) for k in range(3)]
    assert take("Q") == 3
    for b in range(3):
        assert take("Q") == 1
        assert list(np.atleast_1d(take("%df" % take("Q")))) == list(chi[b * 2:(b + 1) * 2])
    assert off == len(data)


def test_thermal_samples_pickle_matches_bincode_fields():
    import pickle
    R, keep, (e, q, chi, smp) = _fake_result()
    betas = np.array([0.5, 1.0, 2.0], np.float32)
    path = tempfile.mktemp(suffix=".pkl")
    _lib.check(_lib.lib().tamc_write_thermal_samples_pickle(C.byref(R), betas.ctypes.data_as(_lib._f32p), 2,
                                                            path.encode()))
    with open(path, "rb") as f:
        obj = pickle.load(f)  # written by this library
    os.unlink(path)
    assert list(obj) == ["compression_level", "instance_size", "beta_arr", "samples", "e", "q", "suscept"]
    assert obj["compression_level"] == 2 and obj["instance_size"] == 10
    assert obj["beta_arr"] == [float(b) for b in betas]
    assert len(obj["samples"]) == 1 and len(obj["samples"][0]) == 4 and obj["samples"][0][1] == list(smp[2:4])
    assert obj["e"] == [[float(x) for x in e[k * 4:(k + 1) * 4]] for k in range(3)]
    assert obj["q"] == [[int(x) for x in q[k * 2:(k + 1) * 2]] for k in range(3)]
    assert obj["suscept"] == [[[float(x) for x in chi[b * 2:(b + 1) * 2]]] for b in range(3)]


@pytest.mark.parametrize("kind", ["rr3", "qubo", "2d"])
def test_instance_create_from_csr_equals_file_build(kind):
    # tamc_instance_create from the fields of a BqmIsingInstance (src/ising.rs:257-263) gives the same
    # coupling_vecs, bias, offset and ICM graph as the file build of the same instance
    rng = np.random.default_rng(12)
    qubo = kind == "qubo"
    i, j, K, n = {"rr3": lambda: W.random_regular3(60, rng, bias_sigma=0.3),
                  "qubo": lambda: W.qubo_sparse(30, 0.2, rng),
                  "2d": lambda: W.lattice_2d(6, rng)}[kind]()
    a = api.BqmIsingInstance.from_lines(i, j, K, qubo=qubo)
    row_ptr, col, J, bias = a.csr()
    w = rng.standard_normal((2, n))
    b = api.BqmIsingInstance.from_csr(row_ptr, col, J, bias=bias, offset=a.offset, suscept=w, qubo=qubo)
    for x, y in zip(a.csr(), b.csr()):
        assert np.array_equal(x, y)
    for x, y in zip(a.graph(), b.graph()):
        assert np.array_equal(x, y)
    assert (b.n, b.nnz, b.offset, b.qubo, b.graph_nodes) == (a.n, a.nnz, a.offset, a.qubo, a.graph_nodes)
    info = _lib.InstanceInfo()
    _lib.check(_lib.lib().tamc_instance_get_info(b._h, C.byref(info)))
    assert info.nchi == 2


def test_instance_create_zero_bias_and_rejections():
    # new_zero_bias (src/ising.rs:266-285): bias None -> zeros, offset 0
    row_ptr = np.array([0, 2, 3, 4], np.uint32)
    col = np.array([1, 2, 0, 0], np.uint32)
    J = np.array([1.0, -2.0, 1.0, -2.0], np.float32)
    g = api.BqmIsingInstance.from_csr(row_ptr, col, J)
    assert np.array_equal(g.csr()[3], np.zeros(3, np.float32)) and g.offset == 0.0
    bad = [
        (np.array([0, 2, 3, 3], np.uint32), col[:3], J[:3]),            # (0,2) without (2,0): not symmetric
        (row_ptr, np.array([1, 2, 0, 5], np.uint32), J),                # column out of range
        (np.array([0, 3, 2, 4], np.uint32), col, J),                    # row_ptr decreasing
        (np.array([0, 2, 3, 5], np.uint32), col, J),                    # row_ptr[n] != nnz
        (np.array([0, 2, 3, 4], np.uint32), np.array([1, 1, 0, 0], np.uint32), J),  # a column twice
    ]
    for rp, c, j in bad:
        with pytest.raises(_lib.TamcError) as e:
            api.BqmIsingInstance.from_csr(rp, c, j)
        assert e.value.status == _lib.TAMC_E_INVALID_ARG
