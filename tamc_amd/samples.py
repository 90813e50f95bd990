"""Readers for the sample files the `tamc` command line and the C ABI writers produce, for analysis
scripts (the files themselves follow the reference's layouts, `src/util.rs:26-36`).

* `read_thermal_samples`: `PtIcmThermalSamples` (`src/pt.rs:116-125`), bincode (legacy encoding:
  little-endian, u64 lengths) or, for `.pkl` files, the pickle the writer emits.
* `read_anneal_state`: `AnnealState` (`src/sa.rs:67-72`), bincode or pickle.
* `read_gla_results`: `GlaResults` (`src/gla.rs:62-70`), bincode or pickle.
* `unpack_states`: bit-packed states (`src/ising.rs:96-121`: bit 1 <=> spin -1, LSB first,
  ceil(N/8) bytes per state) to +-1 arrays.
"""
import os
import pickle
import struct

import numpy as np


class _Reader:
    def __init__(self, data):
        self.data = data
        self.off = 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.data, self.off)
        self.off += struct.calcsize("<" + fmt)
        return v

    def u8(self):
        return self.take("B")[0]

    def u32(self):
        return self.take("I")[0]

    def u64(self):
        return self.take("Q")[0]

    def array(self, dtype):
        n = self.u64()
        dt = np.dtype(dtype).newbyteorder("<")
        a = np.frombuffer(self.data, dtype=dt, count=n, offset=self.off).astype(dtype)
        self.off += n * dt.itemsize
        return a

    def vec(self, item):
        return [item() for _ in range(self.u64())]

    def done(self):
        if self.off != len(self.data):
            raise ValueError("%d trailing bytes after the record" % (len(self.data) - self.off))


def _is_pickle(path):
    return os.path.splitext(str(path))[1] == ".pkl"


def read_thermal_samples(path):
    """PtIcmThermalSamples as a dict: compression_level, instance_size, beta_arr f32[B],
    samples [slot][t*C+c] -> u8[ceil(N/8)], e [b] -> f32[t*C+c], q [b] -> i32[t*P+p],
    suscept [b][k] -> f32[t*P+p]."""
    if _is_pickle(path):
        with open(path, "rb") as f:
            obj = pickle.load(f)  # pickle files are trusted input here, as for any analysis script
        return {
            "compression_level": int(obj["compression_level"]),
            "instance_size": int(obj["instance_size"]),
            "beta_arr": np.asarray(obj["beta_arr"], np.float32),
            "samples": [[np.asarray(s, np.uint8) for s in row] for row in obj["samples"]],
            "e": [np.asarray(v, np.float32) for v in obj["e"]],
            "q": [np.asarray(v, np.int32) for v in obj["q"]],
            "suscept": [[np.asarray(v, np.float32) for v in row] for row in obj["suscept"]],
        }
    with open(path, "rb") as f:
        r = _Reader(f.read())
    out = {"compression_level": r.u8(), "instance_size": r.u64()}
    out["beta_arr"] = r.array(np.float32)
    out["samples"] = r.vec(lambda: r.vec(lambda: r.array(np.uint8)))
    out["e"] = r.vec(lambda: r.array(np.float32))
    out["q"] = r.vec(lambda: r.array(np.int32))
    out["suscept"] = r.vec(lambda: r.vec(lambda: r.array(np.float32)))
    r.done()
    return out


def read_anneal_state(path):
    """AnnealState as a dict: instance_size, num_replicas, bytes (u8, num_replicas * ceil(N/8))."""
    if _is_pickle(path):
        with open(path, "rb") as f:
            obj = pickle.load(f)
        return {"instance_size": int(obj["instance_size"]), "num_replicas": int(obj["num_replicas"]),
                "bytes": np.asarray(obj["bytes"], np.uint8)}
    with open(path, "rb") as f:
        r = _Reader(f.read())
    out = {"instance_size": r.u32(), "num_replicas": r.u32()}
    out["bytes"] = r.array(np.uint8)
    r.done()
    return out


def read_gla_results(path):
    """GlaResults as a dict: instance_size, num_partitions, partition_energies f32[P],
    boundary_energies [(i, j, f32)], final_state u8[], total_energy f32."""
    if _is_pickle(path):
        with open(path, "rb") as f:
            obj = pickle.load(f)
        return {"instance_size": int(obj["instance_size"]), "num_partitions": int(obj["num_partitions"]),
                "partition_energies": np.asarray(obj["partition_energies"], np.float32),
                "boundary_energies": [(int(a), int(b), np.float32(e)) for a, b, e in obj["boundary_energies"]],
                "final_state": np.asarray(obj["final_state"], np.uint8),
                "total_energy": np.float32(obj["total_energy"])}
    with open(path, "rb") as f:
        r = _Reader(f.read())
    out = {"instance_size": r.u32(), "num_partitions": r.u32()}
    out["partition_energies"] = r.array(np.float32)
    out["boundary_energies"] = r.vec(lambda: (r.u32(), r.u32(), np.float32(r.take("f")[0])))
    out["final_state"] = r.array(np.uint8)
    out["total_energy"] = np.float32(r.take("f")[0])
    r.done()
    return out


def unpack_states(packed, n):
    """Bit-packed states (concatenated, ceil(n/8) bytes each) -> int8 array [count, n] of +-1."""
    packed = np.asarray(packed, np.uint8)
    nb = (n + 7) // 8
    if packed.size % nb:
        raise ValueError("%d bytes is not a whole number of %d-byte states" % (packed.size, nb))
    bits = np.unpackbits(packed.reshape(-1, nb), axis=1, bitorder="little")[:, :n]
    return (1 - 2 * bits.astype(np.int8)).astype(np.int8)
