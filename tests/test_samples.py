"""Sample-file readers on hand-built files in the reference layouts (bincode legacy encoding of
PtIcmThermalSamples, src/pt.rs:116-125, and AnnealState, src/sa.rs:67-72) and on pickles with the
writer's structure (structs as dicts, Vecs as lists)."""
import pickle
import struct

import numpy as np

from tamc_amd import samples


def _vec(fmt, xs):
    return struct.pack("<Q", len(xs)) + struct.pack("<%d%s" % (len(xs), fmt), *xs)


def _thermal_bincode(obj):
    b = struct.pack("<BQ", obj["compression_level"], obj["instance_size"])
    b += _vec("f", obj["beta_arr"])
    b += struct.pack("<Q", len(obj["samples"]))
    for row in obj["samples"]:
        b += struct.pack("<Q", len(row)) + b"".join(_vec("B", s) for s in row)
    b += struct.pack("<Q", len(obj["e"])) + b"".join(_vec("f", v) for v in obj["e"])
    b += struct.pack("<Q", len(obj["q"])) + b"".join(_vec("i", v) for v in obj["q"])
    b += struct.pack("<Q", len(obj["suscept"]))
    for row in obj["suscept"]:
        b += struct.pack("<Q", len(row)) + b"".join(_vec("f", v) for v in row)
    return b


OBJ = {
    "compression_level": 1,
    "instance_size": 10,
    "beta_arr": [0.25, 0.5, 1.0],
    "samples": [[[3, 1], [255, 0]], [[7, 2], [0, 3]]],
    "e": [[-1.5, -2.0], [-3.25, 4.0], [0.0, 1.0]],
    "q": [[10, -4], [2, 0], [-10, 8]],
    "suscept": [[[0.5, 1.5]], [[2.5, -1.0]], [[0.0, 3.0]]],
}


def _check(got):
    assert got["compression_level"] == 1 and got["instance_size"] == 10
    assert np.array_equal(got["beta_arr"], np.array(OBJ["beta_arr"], np.float32))
    assert [[list(s) for s in row] for row in got["samples"]] == OBJ["samples"]
    assert [list(v) for v in got["e"]] == OBJ["e"]
    assert [list(v) for v in got["q"]] == OBJ["q"]
    assert [[list(v) for v in row] for row in got["suscept"]] == OBJ["suscept"]
    assert got["q"][0].dtype == np.int32 and got["e"][0].dtype == np.float32


def test_thermal_samples_bincode(tmp_path):
    p = tmp_path / "s.bin"
    p.write_bytes(_thermal_bincode(OBJ))
    _check(samples.read_thermal_samples(p))


def test_thermal_samples_pickle(tmp_path):
    p = tmp_path / "s.pkl"
    with open(p, "wb") as f:
        pickle.dump(OBJ, f, protocol=3)
    _check(samples.read_thermal_samples(p))


def test_thermal_samples_rejects_trailing_bytes(tmp_path):
    p = tmp_path / "s.bin"
    p.write_bytes(_thermal_bincode(OBJ) + b"\0")
    try:
        samples.read_thermal_samples(p)
    except ValueError:
        return
    raise AssertionError("trailing bytes accepted")


def test_anneal_state_and_unpack(tmp_path):
    n = 11
    states = np.array([[1] * n, [-1] * n, [1, -1] * 5 + [-1]], np.int8)
    packed = np.packbits((states < 0).astype(np.uint8), axis=1, bitorder="little").ravel()
    assert packed.size == 3 * 2
    p = tmp_path / "a.bin"
    p.write_bytes(struct.pack("<IIQ", n, 3, packed.size) + packed.tobytes())
    a = samples.read_anneal_state(p)
    assert a["instance_size"] == n and a["num_replicas"] == 3
    assert np.array_equal(samples.unpack_states(a["bytes"], n), states)
    # bit i of byte i // 8, LSB first, 1 <=> -1 (src/ising.rs:96-121)
    assert list(packed[:2]) == [0, 0] and list(packed[2:4]) == [255, 7] and packed[4] == 0b10101010
