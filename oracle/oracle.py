"""ctypes binding of libtamc_oracle.so (the C++ restatement in tamc_oracle.cpp).

Test infrastructure: the product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtamc_oracle.so")

RNG_XOSHIRO, RNG_FEED, RNG_PHILOX = 0, 1, 2

_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i8p = C.POINTER(C.c_int8)
_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


class PtParams(C.Structure):
    _fields_ = [
        ("num_sweeps", C.c_uint32), ("warmup_fraction", C.c_double), ("num_betas", C.c_uint32),
        ("betas", _f32p), ("lo_beta_idx", C.c_uint32), ("icm", C.c_int32), ("num_chains", C.c_uint32),
        ("threaded", C.c_int32), ("sample_enabled", C.c_int32), ("sample", C.c_uint32),
        ("sample_states_enabled", C.c_int32), ("sample_states", C.c_uint32), ("sample_limiting", C.c_uint32),
        ("rng_mode", C.c_int32), ("seed", C.c_uint64), ("num_streams", C.c_uint32),
        ("stream_words", C.POINTER(_u32p)), ("stream_lens", _u64p), ("init_states", _i8p),
        ("sweep_order", C.c_uint32), ("init_energies", _f32p), ("init_nacc", _u32p), ("init_rt", _u32p),
        ("init_tags", _u8p), ("init_hist", _u32p), ("energy_fixed", C.c_uint32),
    ]


class PtResult(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("B", C.c_uint32), ("C", C.c_uint32), ("T", C.c_uint32), ("meas_init", C.c_uint32),
        ("nchi", C.c_uint32), ("npairs", C.c_uint32), ("n_meas", C.c_uint32),
        ("e", _f32p), ("q", _i32p), ("suscept", _f32p),
        ("n_samp_slots", C.c_uint32), ("n_samp_events", C.c_uint32), ("nbytes", C.c_uint32), ("samples", _u8p),
        ("n_gs", C.c_uint32), ("gs_words", C.c_uint32), ("gs_steps", _u32p), ("gs_energies", _f32p),
        ("gs_states", _u64p), ("acceptance_counts", _u32p), ("final_states", _i8p), ("final_energies", _f32p),
        ("final_ids", _u32p), ("num_acceptances", _u32p), ("round_trips", _u32p), ("tags", _u8p),
        ("diffusion_hist", _u32p), ("icm_moves", C.c_uint64), ("attempts", C.c_uint64), ("error", C.c_int32),
    ]


class SaParams(C.Structure):
    _fields_ = [
        ("num_betas", C.c_uint32), ("betas", _f32p), ("num_replicas", C.c_uint32), ("threads", C.c_uint32),
        ("rng_mode", C.c_int32), ("seed", C.c_uint64), ("init_states", _i8p),
    ]


class SaResult(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("R", C.c_uint32), ("min_energy", C.c_float), ("energies", _f32p),
        ("final_states", _i8p), ("attempts", C.c_uint64), ("error", C.c_int32),
    ]


def build():
    """Compile libtamc_oracle.so with its Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    sig = {
        "tamco_philox4x32_10": (None, [_u32p, _u32p, _u32p]),
        "tamco_xoshiro_seed": (None, [C.c_uint64, _u64p]),
        "tamco_xoshiro_next": (C.c_uint64, [_u64p]),
        "tamco_xoshiro_jump": (None, [_u64p]),
        "tamco_gen_index": (C.c_uint32, [C.c_uint32, _u32p, C.c_uint64, _u64p]),
        "tamco_instance_from_file": (C.c_void_p, [C.c_char_p, C.c_int]),
        "tamco_instance_from_lines": (C.c_void_p, [C.c_uint64, _u32p, _u32p, _f64p, C.c_int]),
        "tamco_instance_free": (None, [C.c_void_p]),
        "tamco_instance_n": (C.c_uint32, [C.c_void_p]),
        "tamco_instance_nnz": (C.c_uint32, [C.c_void_p]),
        "tamco_instance_offset": (C.c_float, [C.c_void_p]),
        "tamco_instance_csr": (None, [C.c_void_p, _u32p, _u32p, _f32p, _f32p]),
        "tamco_instance_graph_nnz": (C.c_uint32, [C.c_void_p]),
        "tamco_instance_graph_nodes": (C.c_uint32, [C.c_void_p]),
        "tamco_instance_graph": (None, [C.c_void_p, _u32p, _u32p]),
        "tamco_instance_set_suscept": (None, [C.c_void_p, C.c_uint32, _f64p]),
        "tamco_energy_ref": (C.c_float, [C.c_void_p, _i8p]),
        "tamco_delta_energy": (C.c_float, [C.c_void_p, _i8p, C.c_uint32]),
        "tamco_sweep_feed": (C.c_float, [C.c_void_p, _i8p, C.c_float, _u32p, C.c_uint64, _u64p]),
        "tamco_sweep_philox": (C.c_float, [C.c_void_p, _i8p, C.c_float, C.c_uint64, C.c_uint32, C.c_uint32]),
        "tamco_overlap": (C.c_int64, [_i8p, _i8p, C.c_uint32]),
        "tamco_suscept": (C.c_double, [C.c_void_p, C.c_uint32, _i8p]),
        "tamco_pack_bytes": (None, [_i8p, C.c_uint32, _u8p]),
        "tamco_pack_u64": (None, [_i8p, C.c_uint32, _u64p]),
        "tamco_houdayer_k": (C.c_int64, [C.c_void_p, _i8p, _i8p, C.c_uint32, _u8p]),
        "tamco_num_negative_overlap": (C.c_uint32, [_i8p, _i8p, C.c_uint32]),
        "tamco_exchange": (None, [C.c_uint32, _f32p, _f32p, _u32p, C.c_int, _u32p, C.c_uint64, _u64p, _u8p, _u32p]),
        "tamco_beta_geometric": (None, [C.c_float, C.c_float, C.c_uint32, _f32p]),
        "tamco_chain_step": (None, [C.c_uint32, _u8p, _u8p, _u32p, _u32p, _u32p]),
        "tamco_expf": (C.c_float, [C.c_float]),
        "tamco_pt_run": (C.POINTER(PtResult), [C.c_void_p, C.POINTER(PtParams)]),
        "tamco_pt_result_free": (None, [C.POINTER(PtResult)]),
        "tamco_sa_run": (C.POINTER(SaResult), [C.c_void_p, C.POINTER(SaParams)]),
        "tamco_sa_result_free": (None, [C.POINTER(SaResult)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a, t):
    return a.ctypes.data_as(t)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib.tamco_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return o


class Xoshiro:
    """Xoshiro256++ with seed_from_u64 (splitmix64) and jump, as the reference's RNG."""

    def __init__(self, seed=None, state=None):
        self.s = np.zeros(4, np.uint64)
        if state is not None:
            self.s[:] = state
        else:
            lib.tamco_xoshiro_seed(C.c_uint64(seed), _p(self.s, _u64p))

    def next_u64(self):
        return int(lib.tamco_xoshiro_next(_p(self.s, _u64p)))

    def next_u32(self):
        return self.next_u64() >> 32

    def words(self, count):
        return np.array([self.next_u32() for _ in range(count)], dtype=np.uint32)

    def jump(self):
        lib.tamco_xoshiro_jump(_p(self.s, _u64p))

    def clone(self):
        return Xoshiro(state=self.s.copy())


class ChainTags:
    """Round-trip bookkeeping of one PT chain (tags, acceptance counts, diffusion histogram, round
    trips) driven by given swap decisions; starts as PTState::new (tamc-core/src/pt.rs:27-38)."""

    def __init__(self, B):
        self.B = B
        self.tags = np.zeros(B, np.uint8)
        self.tags[0], self.tags[-1] = 1, 2
        self.nacc = np.zeros(B, np.uint32)
        self.hist = np.zeros((B, 2), np.uint32)
        self.rt = C.c_uint32(0)

    def step(self, swaps):
        sw = np.ascontiguousarray(swaps, np.uint8)
        lib.tamco_chain_step(self.B, _p(sw, _u8p), _p(self.tags, _u8p), _p(self.nacc, _u32p), _p(self.hist, _u32p),
                             C.byref(self.rt))


def beta_geometric(bmin, bmax, nb):
    out = np.zeros(nb, np.float32)
    lib.tamco_beta_geometric(C.c_float(bmin), C.c_float(bmax), nb, _p(out, _f32p))
    return out


def expf(x):
    return lib.tamco_expf(C.c_float(x))


class Instance:
    """BqmIsingInstance restatement (src/ising.rs:257-324)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)
        self.n = lib.tamco_instance_n(self.h)
        nnz = lib.tamco_instance_nnz(self.h)
        self.row_ptr = np.zeros(self.n + 1, np.uint32)
        self.col = np.zeros(max(nnz, 1), np.uint32)
        self.J = np.zeros(max(nnz, 1), np.float32)
        self.bias = np.zeros(max(self.n, 1), np.float32)
        lib.tamco_instance_csr(self.h, _p(self.row_ptr, _u32p), _p(self.col, _u32p), _p(self.J, _f32p),
                               _p(self.bias, _f32p))
        self.col, self.J, self.bias = self.col[:nnz], self.J[:nnz], self.bias[:self.n]
        self.offset = lib.tamco_instance_offset(self.h)
        gnnz = lib.tamco_instance_graph_nnz(self.h)
        self.g_nodes = lib.tamco_instance_graph_nodes(self.h)
        self.g_row_ptr = np.zeros(self.g_nodes + 1, np.uint32)
        self.g_col = np.zeros(max(gnnz, 1), np.uint32)
        lib.tamco_instance_graph(self.h, _p(self.g_row_ptr, _u32p), _p(self.g_col, _u32p))
        self.g_col = self.g_col[:gnnz]
        self.nchi = 0

    @classmethod
    def from_file(cls, path, qubo=False):
        h = lib.tamco_instance_from_file(path.encode(), int(qubo))
        if not h:
            raise IOError(path)
        return cls(h)

    @classmethod
    def from_lines(cls, i, j, K, qubo=False):
        i = np.ascontiguousarray(i, np.uint32)
        j = np.ascontiguousarray(j, np.uint32)
        K = np.ascontiguousarray(K, np.float64)
        return cls(lib.tamco_instance_from_lines(len(i), _p(i, _u32p), _p(j, _u32p), _p(K, _f64p), int(qubo)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.tamco_instance_free(self.h)
            self.h = None

    def set_suscept(self, w):
        w = np.ascontiguousarray(w, np.float64).reshape(-1, self.n)
        self.nchi = w.shape[0]
        lib.tamco_instance_set_suscept(self.h, w.shape[0], _p(w, _f64p))

    def energy(self, s):
        s = np.ascontiguousarray(s, np.int8)
        return lib.tamco_energy_ref(self.h, _p(s, _i8p))

    def delta_energy(self, s, i):
        s = np.ascontiguousarray(s, np.int8)
        return lib.tamco_delta_energy(self.h, _p(s, _i8p), i)

    def sweep_feed(self, s, beta, words, cursor=0):
        """In-place sweep of int8 array s; returns (accepted dE sum, new cursor)."""
        w = np.ascontiguousarray(words, np.uint32)
        cur = C.c_uint64(cursor)
        de = lib.tamco_sweep_feed(self.h, _p(s, _i8p), C.c_float(beta), _p(w, _u32p), len(w), C.byref(cur))
        return de, cur.value

    def sweep_philox(self, s, beta, seed, t, rho):
        return lib.tamco_sweep_philox(self.h, _p(s, _i8p), C.c_float(beta), C.c_uint64(seed), t, rho)

    def houdayer_k(self, s1, s2, k):
        cl = np.zeros(self.n, np.uint8)
        r = lib.tamco_houdayer_k(self.h, _p(s1, _i8p), _p(s2, _i8p), k, _p(cl, _u8p))
        return r, cl.astype(bool)

    def suscept(self, k, ov):
        ov = np.ascontiguousarray(ov, np.int8)
        return lib.tamco_suscept(self.h, k, _p(ov, _i8p))


def overlap(a, b):
    a = np.ascontiguousarray(a, np.int8)
    b = np.ascontiguousarray(b, np.int8)
    return int(lib.tamco_overlap(_p(a, _i8p), _p(b, _i8p), len(a)))


def num_negative_overlap(s1, s2):
    return lib.tamco_num_negative_overlap(_p(s1, _i8p), _p(s2, _i8p), len(s1))


def pack_bytes(s):
    s = np.ascontiguousarray(s, np.int8)
    out = np.zeros((len(s) + 7) // 8, np.uint8)
    lib.tamco_pack_bytes(_p(s, _i8p), len(s), _p(out, _u8p))
    return out


def pack_u64(s):
    s = np.ascontiguousarray(s, np.int8)
    out = np.zeros((len(s) + 63) // 64, np.uint64)
    lib.tamco_pack_u64(_p(s, _i8p), len(s), _p(out, _u64p))
    return out


def gen_index(length, words, cursor=0):
    w = np.ascontiguousarray(words, np.uint32)
    cur = C.c_uint64(cursor)
    r = lib.tamco_gen_index(length, _p(w, _u32p), len(w), C.byref(cur))
    return r, cur.value


def exchange(betas, energies, perm, variant, words, cursor=0):
    betas = np.ascontiguousarray(betas, np.float32)
    energies = np.ascontiguousarray(energies, np.float32)
    perm = np.ascontiguousarray(perm, np.uint32).copy()
    w = np.ascontiguousarray(words, np.uint32)
    B = len(betas)
    acc = np.zeros(max(B - 1, 1), np.uint8)
    nacc = np.zeros(B, np.uint32)
    cur = C.c_uint64(cursor)
    lib.tamco_exchange(B, _p(betas, _f32p), _p(energies, _f32p), _p(perm, _u32p), variant, _p(w, _u32p), len(w),
                       C.byref(cur), _p(acc, _u8p), _p(nacc, _u32p))
    return perm, acc[:B - 1].astype(bool), nacc, cur.value


def _arr(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).copy().astype(dtype)


def pt_run(inst, betas, num_sweeps, num_chains=2, warmup_fraction=0.5, lo_beta_idx=0, icm=True, threaded=False,
           sample=None, sample_states=None, sample_limiting=0, rng_mode=RNG_PHILOX, seed=0, streams=None,
           init_states=None, sweep_order="colour", init_state=None, energy_fixed=None):
    """Run PT-ICM with the restated loop; returns a dict of numpy arrays (slot order).
    init_state: optional dict with the rest of a PTState to resume from (energies, num_acceptances,
    round_trips, tags, diffusion_hist; each optional), in slot order.
    energy_fixed: the production path's exact tracked-energy accounting (default: on in Philox mode,
    which restates the production path; the reference's f32 accumulation otherwise)."""
    betas = np.ascontiguousarray(betas, np.float32)
    B = len(betas)
    P = PtParams()
    P.num_sweeps = num_sweeps
    P.warmup_fraction = warmup_fraction
    P.num_betas = B
    P.betas = _p(betas, _f32p)
    P.lo_beta_idx = lo_beta_idx
    P.icm = int(icm)
    P.num_chains = num_chains
    P.threaded = int(threaded)
    P.sample_enabled = int(sample is not None)
    P.sample = sample or 0
    P.sample_states_enabled = int(sample_states is not None)
    P.sample_states = sample_states or 0
    P.sample_limiting = sample_limiting
    P.rng_mode = rng_mode
    P.seed = seed
    P.sweep_order = {"colour": 0, "sequential": 1}[sweep_order]
    P.energy_fixed = int(rng_mode == RNG_PHILOX if energy_fixed is None else energy_fixed)
    keep = []
    if streams is not None:
        arrs = [np.ascontiguousarray(s, np.uint32) for s in streams]
        keep.append(arrs)
        ptrs = (_u32p * len(arrs))(*[_p(a, _u32p) for a in arrs])
        lens = np.array([len(a) for a in arrs], np.uint64)
        keep += [ptrs, lens]
        P.num_streams = len(arrs)
        P.stream_words = ptrs
        P.stream_lens = _p(lens, _u64p)
    if init_states is not None:
        init = np.ascontiguousarray(init_states, np.int8)
        keep.append(init)
        P.init_states = _p(init, _i8p)
    for key, field, dt, pt in [("energies", "init_energies", np.float32, _f32p),
                               ("num_acceptances", "init_nacc", np.uint32, _u32p),
                               ("round_trips", "init_rt", np.uint32, _u32p), ("tags", "init_tags", np.uint8, _u8p),
                               ("diffusion_hist", "init_hist", np.uint32, _u32p)]:
        if init_state is not None and init_state.get(key) is not None:
            a = np.ascontiguousarray(init_state[key], dt)
            keep.append(a)
            setattr(P, field, _p(a, pt))
    rp = lib.tamco_pt_run(inst.h, C.byref(P))
    R = rp.contents
    try:
        if R.error == 3:
            raise ValueError("invalid PT parameters")
        n, B, Cc, npairs, nm = R.n, R.B, R.C, R.npairs, R.n_meas
        out = dict(error=R.error, n=n, B=B, C=Cc, T=R.T, meas_init=R.meas_init, n_meas=nm,
                   e=_arr(R.e, B * nm * Cc, np.float32).reshape(B, nm * Cc),
                   q=_arr(R.q, B * nm * npairs, np.int32).reshape(B, nm * npairs),
                   suscept=_arr(R.suscept, B * R.nchi * nm * npairs, np.float32).reshape(B, R.nchi, nm * npairs),
                   samples=_arr(R.samples, R.n_samp_slots * R.n_samp_events * Cc * R.nbytes, np.uint8).reshape(
                       R.n_samp_slots, R.n_samp_events * Cc, R.nbytes),
                   gs_steps=_arr(R.gs_steps, R.n_gs, np.uint32),
                   gs_energies=_arr(R.gs_energies, R.n_gs, np.float32),
                   gs_states=_arr(R.gs_states, R.n_gs * R.gs_words, np.uint64).reshape(R.n_gs, R.gs_words),
                   acceptance_counts=_arr(R.acceptance_counts, B, np.uint32),
                   final_states=_arr(R.final_states, Cc * B * n, np.int8).reshape(Cc, B, n),
                   final_energies=_arr(R.final_energies, Cc * B, np.float32).reshape(Cc, B),
                   final_ids=_arr(R.final_ids, Cc * B, np.uint32).reshape(Cc, B),
                   num_acceptances=_arr(R.num_acceptances, Cc * B, np.uint32).reshape(Cc, B),
                   round_trips=_arr(R.round_trips, Cc, np.uint32),
                   tags=_arr(R.tags, Cc * B, np.uint8).reshape(Cc, B),
                   diffusion_hist=_arr(R.diffusion_hist, Cc * B * 2, np.uint32).reshape(Cc, B, 2),
                   icm_moves=R.icm_moves, attempts=R.attempts)
    finally:
        lib.tamco_pt_result_free(rp)
    return out


def sa_run(inst, betas, num_replicas, threads=1, rng_mode=RNG_PHILOX, seed=0, init_states=None):
    betas = np.ascontiguousarray(betas, np.float32)
    P = SaParams()
    P.num_betas = len(betas)
    P.betas = _p(betas, _f32p)
    P.num_replicas = num_replicas
    P.threads = threads
    P.rng_mode = rng_mode
    P.seed = seed
    keep = None
    if init_states is not None:
        keep = np.ascontiguousarray(init_states, np.int8)
        P.init_states = _p(keep, _i8p)
    rp = lib.tamco_sa_run(inst.h, C.byref(P))
    R = rp.contents
    try:
        if R.error:
            raise ValueError("invalid SA parameters")
        out = dict(n=R.n, R=R.R, min_energy=R.min_energy, energies=_arr(R.energies, R.R, np.float32),
                   final_states=_arr(R.final_states, R.R * R.n, np.int8).reshape(R.R, R.n), attempts=R.attempts)
    finally:
        lib.tamco_sa_result_free(rp)
    return out
