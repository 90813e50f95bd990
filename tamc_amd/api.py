"""Python mirror of the reference's runner API over the C ABI.

Names follow the reference: ``BqmIsingInstance`` (src/ising.rs:257-324), ``BetaOptions`` and
``PtIcmParams`` (src/pt.rs:229-290), ``PtIcmRunner`` (src/pt.rs:298-593), ``SaParams`` /
``SaRunner`` (src/sa.rs:21-211).  Results are dictionaries of numpy arrays in the reference's
layouts (``e[b][t*C + c]``, ``q[b][t*npairs + p]``, ...).  Every computation runs in
libtamc_gpu.so on the GPU; nothing here computes a result itself.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from ._lib import (BatchOptions, BatchStats, InstanceDesc, InstanceInfo, PtParams as _CPt, PtResult, PtState,
                   RequestTimes, SaParams as _CSa,
                   SaResult, ServerOptions, ServerStats, TAMC_E_PENDING, check, lib, ptr, _f32p, _f64p, _i8p, _i64p,
                   _u32p, _u64p, _u8p)
from . import _lib


class BqmIsingInstance:
    """Ising/QUBO instance held by the library (BqmIsingInstance, src/ising.rs:257-324)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        info = InstanceInfo()
        check(lib().tamc_instance_get_info(self._h, C.byref(info)))
        self.n, self.nnz, self.offset, self.qubo = info.n, info.nnz, info.offset, bool(info.qubo)
        self.graph_nodes = info.graph_nodes

    @classmethod
    def from_instance_file(cls, path, qubo=False):
        h = C.c_void_p()
        check(lib().tamc_instance_from_file(path.encode(), int(qubo), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_lines(cls, i, j, K, qubo=False):
        i = np.ascontiguousarray(i, np.uint32)
        j = np.ascontiguousarray(j, np.uint32)
        K = np.ascontiguousarray(K, np.float64)
        h = C.c_void_p()
        check(lib().tamc_instance_from_lines(len(i), ptr(i, _u32p), ptr(j, _u32p), ptr(K, _f64p), int(qubo),
                                             C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr(cls, row_ptr, col, J, bias=None, offset=0.0, suscept=None, qubo=False):
        """An instance from the fields of BqmIsingInstance (src/ising.rs:257-263): coupling_vecs as
        CSR (row order kept, QUBO self entries included), bias (None: zero, as new_zero_bias),
        offset and suscept_coefs."""
        rp = np.ascontiguousarray(row_ptr, np.uint32)
        cl = np.ascontiguousarray(col, np.uint32)
        Jv = np.ascontiguousarray(J, np.float32)
        n = len(rp) - 1
        keep = [rp, cl, Jv]
        d = InstanceDesc()
        d.n, d.nnz = n, len(cl)
        d.row_ptr, d.col, d.J = ptr(rp, _u32p), ptr(cl, _u32p), ptr(Jv, _f32p)
        if bias is not None:
            b = np.ascontiguousarray(bias, np.float32)
            keep.append(b)
            d.bias = ptr(b, _f32p)
        d.offset = offset
        if suscept is not None:
            w = np.ascontiguousarray(suscept, np.float64).reshape(-1, n)
            keep.append(w)
            d.nchi, d.suscept = w.shape[0], ptr(w, _f64p)
        d.qubo = int(qubo)
        h = C.c_void_p()
        check(lib().tamc_instance_create(C.byref(d), C.byref(h)))
        return cls(h.value)

    def with_suscept(self, w):
        w = np.ascontiguousarray(w, np.float64).reshape(-1, self.n)
        check(lib().tamc_instance_set_suscept(self._h, w.shape[0], ptr(w, _f64p)))
        return self

    def add_suscept_file(self, path):
        check(lib().tamc_instance_add_suscept_file(self._h, path.encode()))

    def size(self):
        return self.n

    def csr(self):
        row_ptr = np.zeros(self.n + 1, np.uint32)
        col = np.zeros(max(self.nnz, 1), np.uint32)
        J = np.zeros(max(self.nnz, 1), np.float32)
        bias = np.zeros(max(self.n, 1), np.float32)
        check(lib().tamc_instance_get_csr(self._h, ptr(row_ptr, _u32p), ptr(col, _u32p), ptr(J, _f32p),
                                          ptr(bias, _f32p)))
        return row_ptr, col[:self.nnz], J[:self.nnz], bias[:self.n]

    def graph(self):
        """ICM graph (coupling CsMat pattern) as (row_ptr, col)."""
        info = InstanceInfo()
        check(lib().tamc_instance_get_info(self._h, C.byref(info)))
        row_ptr = np.zeros(info.graph_nodes + 1, np.uint32)
        col = np.zeros(max(info.graph_nnz, 1), np.uint32)
        check(lib().tamc_instance_get_graph(self._h, ptr(row_ptr, _u32p), ptr(col, _u32p)))
        return row_ptr, col[:info.graph_nnz]

    def induced_subgraph(self, nodes):
        """induced_subgraph_instance_sorted (src/ising.rs:327-355) on ascending site indices."""
        nodes = np.ascontiguousarray(nodes, np.uint32)
        h = C.c_void_p()
        check(lib().tamc_instance_induced_subgraph(self._h, ptr(nodes, _u32p), len(nodes), C.byref(h)))
        return BqmIsingInstance(h.value)

    # -- parity surface ------------------------------------------------------------------
    def energy(self, states):
        s = np.ascontiguousarray(np.atleast_2d(states), np.int8)
        out = np.zeros(s.shape[0], np.float32)
        check(lib().tamc_energy(self._h, s.shape[0], ptr(s, _i8p), ptr(out, _f32p)))
        return out

    def delta_energy(self, states):
        s = np.ascontiguousarray(np.atleast_2d(states), np.int8)
        out = np.zeros(s.shape, np.float32)
        check(lib().tamc_delta_energy(self._h, s.shape[0], ptr(s, _i8p), ptr(out, _f32p)))
        return out

    def houdayer(self, s1, s2, k):
        """Houdayer move on state pairs with seeds k (index into the ascending q<0 list)."""
        s1 = np.ascontiguousarray(np.atleast_2d(s1), np.int8).copy()
        s2 = np.ascontiguousarray(np.atleast_2d(s2), np.int8).copy()
        k = np.ascontiguousarray(np.atleast_1d(k), np.uint32)
        sizes = np.zeros(s1.shape[0], np.int64)
        cl = np.zeros(s1.shape, np.uint8)
        check(lib().tamc_houdayer(self._h, s1.shape[0], ptr(s1, _i8p), ptr(s2, _i8p), ptr(k, _u32p),
                                  ptr(sizes, _i64p), ptr(cl, _u8p)))
        return s1, s2, sizes, cl.astype(bool)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().tamc_instance_destroy(h)
            except Exception:
                pass
            self._h = None


def overlap(a, b):
    a = np.ascontiguousarray(np.atleast_2d(a), np.int8)
    b = np.ascontiguousarray(np.atleast_2d(b), np.int8)
    q = np.zeros(a.shape[0], np.int64)
    check(lib().tamc_overlap(a.shape[1], a.shape[0], ptr(a, _i8p), ptr(b, _i8p), ptr(q, _i64p)))
    return q


def pack_bytes(states):
    s = np.ascontiguousarray(np.atleast_2d(states), np.int8)
    out = np.zeros((s.shape[0], (s.shape[1] + 7) // 8), np.uint8)
    check(lib().tamc_pack_bytes(s.shape[1], s.shape[0], ptr(s, _i8p), ptr(out, _u8p)))
    return out


@dataclass
class BetaOptions:
    """BetaOptions (src/pt.rs:229-257): Geometric(beta_min, beta_max, num_beta) or Arr(values)."""
    geometric: bool = True
    beta_min: float = 0.1
    beta_max: float = 10.0
    num_beta: int = 8
    arr: Optional[List[float]] = None

    @classmethod
    def Geometric(cls, beta_min, beta_max, num_beta):
        return cls(True, beta_min, beta_max, num_beta)

    @classmethod
    def Arr(cls, values):
        return cls(False, arr=list(values))


@dataclass
class PtIcmParams:
    """PtIcmParams (src/pt.rs:259-290) plus the production RNG seed."""
    num_sweeps: int = 256
    warmup_fraction: float = 0.5
    beta: BetaOptions = field(default_factory=lambda: BetaOptions.Geometric(0.1, 10.0, 8))
    lo_beta: Optional[float] = None
    lo_num_beta: Optional[int] = None
    icm: bool = True
    num_replica_chains: int = 2
    threads: int = 1
    sample: Optional[int] = 32
    sample_states: Optional[int] = 64
    sample_limiting: Optional[int] = 0
    seed: int = 0
    sweep_order: str = "colour"  # "colour" (production) or "sequential" (the reference's site order)

    def resolve(self):
        """PtIcmRunner::new (src/pt.rs:307-351): (betas f32 array, lo_beta_idx)."""
        b = self.beta
        nb_cap = b.num_beta if b.geometric else len(b.arr)
        out = np.zeros(max(nb_cap, 1), np.float32)
        arr = np.ascontiguousarray(b.arr if b.arr is not None else [0.0], np.float32)
        nb = C.c_uint32()
        lo = C.c_uint32()
        check(lib().tamc_pt_resolve(int(b.geometric), C.c_float(b.beta_min), C.c_float(b.beta_max), b.num_beta,
                                    ptr(arr, _f32p), len(b.arr) if b.arr is not None else 0,
                                    C.c_float(float("nan") if self.lo_beta is None else self.lo_beta),
                                    -1 if self.lo_num_beta is None else int(self.lo_num_beta), int(self.icm),
                                    ptr(out, _f32p), C.byref(nb), C.byref(lo)))
        return out[:nb.value].copy(), lo.value

    def to_c(self, betas, lo):
        P = _CPt()
        P.num_sweeps = self.num_sweeps
        P.warmup_fraction = self.warmup_fraction
        P.num_betas = len(betas)
        P.betas = ptr(betas, _f32p)
        P.lo_beta_idx = lo
        P.icm = int(self.icm)
        P.num_replica_chains = self.num_replica_chains
        P.threads = self.threads
        P.sample_enabled = int(self.sample is not None)
        P.sample = self.sample or 0
        P.sample_states_enabled = int(self.sample_states is not None)
        P.sample_states = self.sample_states or 0
        P.sample_limiting = self.sample_limiting or 0
        P.seed = self.seed
        P.sweep_order = {"colour": 0, "sequential": 1}[self.sweep_order]
        return P


def _arr(p, count, dtype):
    # one copy of count elements out of library memory (string_at + frombuffer is several times
    # cheaper per array than numpy.ctypeslib.as_array + copy, which matters for large batches)
    dt = np.dtype(dtype)
    if count == 0:
        return np.zeros(0, dt)
    return np.frombuffer(bytearray(C.string_at(p, count * dt.itemsize)), dt)


def pt_result_to_dict(R):
    n, B, Cc, npairs, nm = R.n, R.B, R.C, R.npairs, R.n_meas
    return dict(
        n=n, B=B, C=Cc, T=R.T, meas_init=R.meas_init, n_meas=nm,
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
        icm_moves=R.icm_moves, attempts=R.attempts, timing_us=R.timing_us, device_ms=R.device_ms)


class PtIcmRunner:
    """PtIcmRunner (src/pt.rs:298-593) on the GPU."""

    def __init__(self, instance, params):
        self.instance = instance
        self.params = params
        self.betas, self.lo_beta_idx = params.resolve()

    def _c(self):
        return self.params.to_c(self.betas, self.lo_beta_idx)

    def run(self, init_states=None, sample_output=None):
        """Production path; returns the result dictionary.  With sample_output, the
        PtIcmThermalSamples are also written there as bincode (src/pt.rs:640-654)."""
        P = self._c()
        out = C.POINTER(PtResult)()
        init = None if init_states is None else np.ascontiguousarray(init_states, np.int8)
        check(lib().tamc_pt_run(self.instance._h, C.byref(P), None if init is None else ptr(init, _i8p),
                                C.byref(out)))
        try:
            if sample_output is not None:
                check(lib().tamc_write_thermal_samples(out, ptr(self.betas, _f32p), 0, str(sample_output).encode()))
            return pt_result_to_dict(out.contents)
        finally:
            lib().tamc_pt_result_free(out)

    run_parallel = run

    def run_from(self, state):
        """PtIcmRunner::run(Some(states)) (src/pt.rs:397-408): continue from a full PTState, given
        as a previous result (its final_states, final_energies, num_acceptances, round_trips, tags
        and diffusion_hist) or a dict with the keys states, energies, num_acceptances, round_trips,
        tags, diffusion_hist (all but states optional)."""
        alias = {"final_states": "states", "final_energies": "energies"}
        st = {alias.get(k, k): v for k, v in state.items()}
        P = self._c()
        S = PtState()
        keep = []
        for name, dt, pt in [("states", np.int8, _i8p), ("energies", np.float32, _f32p),
                             ("num_acceptances", np.uint32, _u32p), ("round_trips", np.uint32, _u32p),
                             ("tags", np.uint8, _u8p), ("diffusion_hist", np.uint32, _u32p)]:
            if st.get(name) is not None:
                a = np.ascontiguousarray(st[name], dt)
                keep.append(a)
                setattr(S, name, ptr(a, pt))
        out = C.POINTER(PtResult)()
        check(lib().tamc_pt_run_resume(self.instance._h, C.byref(P), C.byref(S), C.byref(out)))
        try:
            return pt_result_to_dict(out.contents)
        finally:
            lib().tamc_pt_result_free(out)

    def run_lockstep(self, init_states, streams):
        """Lock-step mode: every random number from the given u32 word streams (see the C header)."""
        P = self._c()
        init = np.ascontiguousarray(init_states, np.int8)
        arrs = [np.ascontiguousarray(s, np.uint32) for s in streams]
        ptrs = (_u32p * len(arrs))(*[ptr(a, _u32p) for a in arrs])
        lens = np.array([len(a) for a in arrs], np.uint64)
        out = C.POINTER(PtResult)()
        check(lib().tamc_pt_run_lockstep(self.instance._h, C.byref(P), ptr(init, _i8p), len(arrs), ptrs,
                                         ptr(lens, _u64p), C.byref(out)))
        try:
            return pt_result_to_dict(out.contents)
        finally:
            lib().tamc_pt_result_free(out)


def _batch_args(instances, params_list):
    resolved = [p.resolve() for p in params_list]
    cps = (_CPt * len(instances))()
    for k, (p, (b, lo)) in enumerate(zip(params_list, resolved)):
        cps[k] = p.to_c(b, lo)
    hs = (C.c_void_p * len(instances))(*[inst._h.value for inst in instances])
    return resolved, cps, hs


def run_pt_batch(instances, params_list, waves_per_group=0, sweeps_per_launch=0, use_graph=False, plain=False,
                 autotune=False, init_states=None, task_tiling=0, icm_growth=0, teams=0):
    """Many PT-ICM solves: batched persistent kernel (default) or the plain one-instance-per-launch path.
    init_states: optional list (None entries allowed) of [C][B][n] +-1 start states per solve.
    task_tiling: 1 = a level's tasks dealt to the waves in turn, 2 = one contiguous run per wave,
    0 = tuned with `autotune` (else 1).  icm_growth: 1 = strided sites per thread, 2 = compact tiles,
    0 = tuned with `autotune` (else tiles when the instance's rows hold more than 4 couplings, or 4 on
    at most 4096 sites).
    teams: 0 = large groups' sweeps shared by teams of work groups where a cost model (and, with
    `autotune`, a timing) says it pays, 1 = never, K >= 2 = every eligible group as a team of K."""
    resolved, cps, hs = _batch_args(instances, params_list)
    outs = (C.POINTER(PtResult) * len(instances))()
    stats = BatchStats()
    if plain:
        check(lib().tamc_pt_run_plain(len(instances), hs, cps, outs, C.byref(stats)))
    elif init_states is not None:
        keep = [None if s is None else np.ascontiguousarray(s, np.int8) for s in init_states]
        ptrs = (_i8p * len(instances))(*[_i8p() if s is None else ptr(s, _i8p) for s in keep])
        opt = BatchOptions(waves_per_group, sweeps_per_launch, int(use_graph), int(autotune), task_tiling,
                           icm_growth, teams)
        check(lib().tamc_pt_run_batch_init(len(instances), hs, cps, ptrs, C.byref(opt), outs, C.byref(stats)))
    else:
        opt = BatchOptions(waves_per_group, sweeps_per_launch, int(use_graph), int(autotune), task_tiling,
                           icm_growth, teams)
        check(lib().tamc_pt_run_batch(len(instances), hs, cps, C.byref(opt), outs, C.byref(stats)))
    res = []
    try:
        for k in range(len(instances)):
            res.append(pt_result_to_dict(outs[k].contents))
    finally:
        for k in range(len(instances)):
            lib().tamc_pt_result_free(outs[k])
    st = dict(device_ms=stats.device_ms, wall_ms=stats.wall_ms, attempts=stats.attempts,
              waves_per_group=stats.waves_per_group, sweeps_per_launch=stats.sweeps_per_launch,
              launches=stats.launches, groups=stats.groups, phase_ms=list(stats.phase_ms),
              icm_passes=stats.icm_passes, prep_ms=stats.prep_ms, post_ms=stats.post_ms,
              groups_busy_ms=stats.groups_busy_ms, task_tiling=stats.task_tiling, graph=bool(stats.graph),
              icm_growth=stats.icm_growth, teams=stats.teams, work_groups=stats.work_groups)
    return res, st


def pt_opt_update(betas, momentum, dif_hists, round_trips, num_sweeps, num_chains, alpha, m):
    """One temperature-set update (tamc_pt_opt_update).  Returns (new_betas, new_momentum, err, taus,
    updated)."""
    b = np.ascontiguousarray(betas, np.float32)
    mo = np.array(momentum, np.float32)
    h = np.ascontiguousarray(dif_hists, np.uint32)
    r = np.ascontiguousarray(round_trips, np.uint32)
    out = np.zeros(len(b), np.float32)
    taus = np.zeros(len(r), np.float32)
    err = C.c_float()
    upd = C.c_int()
    check(lib().tamc_pt_opt_update(len(b), ptr(b, _f32p), ptr(mo, _f32p), len(r), ptr(h, _u32p), ptr(r, _u32p),
                                   int(num_sweeps), int(num_chains), float(alpha), float(m), ptr(out, _f32p),
                                   C.byref(err), ptr(taus, _f32p), C.byref(upd)))
    return out, mo, np.float32(err.value), taus, bool(upd.value)


def pt_optimize_beta(instances, params, num_iters, alpha, m, tol, log=None):
    """pt_optimize_beta (src/bin/tamc-pt-opt.rs:34-198): every iteration runs PT on all instances
    in one batched launch (iteration i, instance k: seed params.seed + i * count + k) from the previous
    iteration's final states, then updates the temperature set.  Returns (params with the optimised
    Arr betas and doubled num_sweeps, tau history [iterations][instances])."""
    import dataclasses
    params = dataclasses.replace(params)
    count = len(instances)
    betas = params.resolve()[0]
    momentum = np.array(betas, np.float32)
    init = [None] * count
    tau_hist = []
    for it in range(num_iters):
        betas = params.resolve()[0]
        ps = [dataclasses.replace(params, seed=params.seed + it * count + k) for k in range(count)]
        res, _ = run_pt_batch(instances, ps, init_states=init)
        hists = np.stack([r["diffusion_hist"].sum(axis=0, dtype=np.uint64).astype(np.uint32) for r in res])
        rts = np.array([int(r["round_trips"].sum()) for r in res], np.uint32)
        new, momentum_next, err, taus, updated = pt_opt_update(betas, momentum, hists, rts, params.num_sweeps,
                                                               params.num_replica_chains, alpha, m)
        tau_hist.append(taus)
        if log:
            log(it, betas, new, err, taus, updated)
        if not updated:
            continue
        momentum = momentum_next
        params.beta = BetaOptions.Arr([float(x) for x in new])
        init = [r["final_states"] for r in res]
        if err < tol:
            break
    params.num_sweeps *= 2
    return params, np.array(tau_hist, np.float32)


def random_percolation(instance, p, num_samples, seed=0):
    """random_percolation (src/percolation.rs:11-75) for num_samples samples on the GPU: returns
    (max_cluster u32[num_samples], percolates bool[num_samples])."""
    cl = np.zeros(max(num_samples, 1), np.uint32)
    pc = np.zeros(max(num_samples, 1), np.uint8)
    check(lib().tamc_percolate(instance._h, float(p), int(num_samples), int(seed), ptr(cl, _u32p), ptr(pc, _u8p)))
    return cl[:num_samples], pc[:num_samples].astype(bool)


@dataclass
class GlaParams:
    """GlaParams (src/gla.rs:28-33).  num_threads is accepted for fidelity: the sub-solves all run in
    one batched launch."""
    num_threads: int
    pt_params: "PtIcmParams"
    partition_file: str = ""


def _partition_list(P):
    offs = _arr(P.offsets, P.num_partitions + 1, np.uint32)
    sites = _arr(P.sites, int(offs[-1]), np.uint32)
    return [sites[offs[k]:offs[k + 1]].copy() for k in range(P.num_partitions)]


def read_partition(instance, path):
    """GlaParams::read_graph_partition (src/gla.rs:35-59): the partitions of a partition file, each
    sorted ascending."""
    out = C.POINTER(_lib.Partition)()
    check(lib().tamc_partition_from_file(instance._h, str(path).encode(), C.byref(out)))
    try:
        return _partition_list(out.contents)
    finally:
        lib().tamc_partition_free(out)


def run_gla(instance, pt_params, partitions=None, partition_file=None, output_file=None):
    """run_gla (src/gla.rs:122-196) on the GPU: the partitions' sub-instances solved in one batched
    launch (sub-solve k with seed pt_params.seed + k), joined, with boundary energies.  Returns the
    GlaResults fields plus the joined state; writes them to output_file (bincode, pickle for .pkl)
    when given."""
    out = C.POINTER(_lib.Partition)()
    if partition_file is not None:
        check(lib().tamc_partition_from_file(instance._h, str(partition_file).encode(), C.byref(out)))
    else:
        parts = [np.ascontiguousarray(p, np.uint32) for p in partitions]
        offs = np.zeros(len(parts) + 1, np.uint32)
        offs[1:] = np.cumsum([len(p) for p in parts])
        sites = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(1, np.uint32), np.uint32)
        check(lib().tamc_partition_create(instance._h, len(parts), ptr(offs, _u32p), ptr(sites, _u32p),
                                          C.byref(out)))
    try:
        betas, lo = pt_params.resolve()
        P = pt_params.to_c(betas, lo)
        res = C.POINTER(_lib.GlaResult)()
        check(lib().tamc_gla_run(instance._h, out, C.byref(P), C.byref(res)))
    finally:
        parts_list = _partition_list(out.contents)
        lib().tamc_partition_free(out)
    try:
        R = res.contents
        nb = R.num_boundaries
        pairs = _arr(R.boundary_pairs, 2 * nb, np.uint32).reshape(nb, 2)
        bnd = _arr(R.boundary_energies, nb, np.float32)
        d = dict(instance_size=R.instance_size, num_partitions=R.num_partitions,
                 partition_energies=_arr(R.partition_energies, R.num_partitions, np.float32),
                 boundary_energies=[(int(a), int(b), float(e)) for (a, b), e in zip(pairs, bnd)],
                 boundary_energies_f32=bnd, final_state=[], total_energy=np.float32(R.total_energy),
                 bulk_energy=np.float32(R.bulk_energy), boundary_energy=np.float32(R.boundary_energy),
                 joined_state=_arr(R.joined_state, R.instance_size, np.int8), partitions=parts_list,
                 attempts=R.attempts, device_ms=R.device_ms)
        if output_file is not None:
            check(lib().tamc_write_gla_results(res, str(output_file).encode()))
        return d
    finally:
        lib().tamc_gla_result_free(res)


class Server:
    """Continuous-batching PT-ICM and SA request loop (tamc_server_* in the C header).

    submit() queues a solve and returns a ticket; poll() returns (result, times) once it has
    finished, or None while it runs.  Instances must outlive their requests."""

    def __init__(self, pool_bytes=0, max_wave_slots=0, slice_ms=0.0, waves_per_group=0):
        self._h = C.c_void_p()
        opt = ServerOptions(pool_bytes, max_wave_slots, slice_ms, waves_per_group)
        check(lib().tamc_server_create(C.byref(opt), C.byref(self._h)))
        self._keep = {}

    def submit(self, instance, params, init_states=None):
        betas, lo = params.resolve()
        P = params.to_c(betas, lo)
        init = None if init_states is None else np.ascontiguousarray(init_states, np.int8)
        t = C.c_uint64()
        check(lib().tamc_server_submit(self._h, instance._h, C.byref(P), None if init is None else ptr(init, _i8p),
                                       C.byref(t)))
        self._keep[t.value] = instance
        return t.value

    def poll(self, ticket, wait_ms=0):
        out = C.POINTER(PtResult)()
        tm = RequestTimes()
        st = lib().tamc_server_poll(self._h, ticket, int(wait_ms), C.byref(out), C.byref(tm))
        if st == TAMC_E_PENDING:
            return None
        self._keep.pop(ticket, None)
        check(st)
        try:
            res = pt_result_to_dict(out.contents)
        finally:
            lib().tamc_pt_result_free(out)
        return res, dict(queued_ms=tm.queued_ms, total_ms=tm.total_ms, slices=tm.slices)

    def wait(self, ticket):
        return self.poll(ticket, wait_ms=-1)

    def submit_sa(self, instance, params):
        """Queue an SA solve (SaRunner::run, src/sa.rs:110-165); returns its ticket."""
        betas = params.betas()
        P = _CSa(len(betas), ptr(betas, _f32p), params.num_replicas, params.threads, params.seed)
        t = C.c_uint64()
        check(lib().tamc_server_submit_sa(self._h, instance._h, C.byref(P), C.byref(t)))
        self._keep[t.value] = instance
        return t.value

    def poll_sa(self, ticket, wait_ms=0):
        out = C.POINTER(SaResult)()
        tm = RequestTimes()
        st = lib().tamc_server_poll_sa(self._h, ticket, int(wait_ms), C.byref(out), C.byref(tm))
        if st == TAMC_E_PENDING:
            return None
        self._keep.pop(ticket, None)
        check(st)
        try:
            res = _sa_dict(out.contents)
        finally:
            lib().tamc_sa_result_free(out)
        return res, dict(queued_ms=tm.queued_ms, total_ms=tm.total_ms, slices=tm.slices)

    def wait_sa(self, ticket):
        return self.poll_sa(ticket, wait_ms=-1)

    def cancel(self, ticket):
        check(lib().tamc_server_cancel(self._h, ticket))
        self._keep.pop(ticket, None)

    def stats(self):
        s = ServerStats()
        check(lib().tamc_server_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            lib().tamc_server_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class SaParams:
    """SaParams (src/sa.rs:21-26) plus the production RNG seed."""
    beta: BetaOptions = field(default_factory=lambda: BetaOptions.Geometric(0.1, 10.0, 256))
    num_replicas: int = 1
    threads: int = 1
    seed: int = 0

    def betas(self):
        b = self.beta
        if b.geometric:
            p = PtIcmParams(beta=b, icm=False)
            return p.resolve()[0]
        return np.ascontiguousarray(b.arr, np.float32)


def _sa_dict(R):
    return dict(n=R.n, R=R.R, min_energy=R.min_energy, energies=_arr(R.energies, R.R, np.float32),
                final_states=_arr(R.final_states, R.R * R.n, np.int8).reshape(R.R, R.n), attempts=R.attempts,
                timing_us=R.timing_us, device_ms=R.device_ms)


def run_sa_batch(instances, params_list):
    keep = [p.betas() for p in params_list]
    cps = (_CSa * len(instances))()
    for k, (p, b) in enumerate(zip(params_list, keep)):
        cps[k] = _CSa(len(b), ptr(b, _f32p), p.num_replicas, p.threads, p.seed)
    hs = (C.c_void_p * len(instances))(*[inst._h.value for inst in instances])
    outs = (C.POINTER(SaResult) * len(instances))()
    stats = BatchStats()
    check(lib().tamc_sa_run_batch(len(instances), hs, cps, outs, C.byref(stats)))
    res = []
    try:
        for k in range(len(instances)):
            res.append(_sa_dict(outs[k].contents))
    finally:
        for k in range(len(instances)):
            lib().tamc_sa_result_free(outs[k])
    return res, dict(device_ms=stats.device_ms, wall_ms=stats.wall_ms, attempts=stats.attempts)


class SaRunner:
    """SaRunner (src/sa.rs:90-211) on the GPU."""

    def __init__(self, instance, params):
        self.instance = instance
        self.params = params

    def run(self):
        return run_sa_batch([self.instance], [self.params])[0][0]

    run_parallel = run
