"""ctypes binding of libtamc_gpu.so (the C ABI declared in include/tamc_gpu.h).

The library is loaded from tamc_amd/lib/.  There is no fallback: if it is missing or the device
is not a gfx950, calls fail with TamcError.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libtamc_gpu.so")

(TAMC_OK, TAMC_E_INVALID_ARG, TAMC_E_IO, TAMC_E_DEVICE, TAMC_E_NOMEM, TAMC_E_UNSUPPORTED, TAMC_E_STREAM,
 TAMC_E_PENDING) = range(8)
_STATUS = {1: "INVALID_ARG", 2: "IO", 3: "DEVICE", 4: "NOMEM", 5: "UNSUPPORTED", 6: "STREAM", 7: "PENDING"}

_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i8p = C.POINTER(C.c_int8)
_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class TamcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("tamc %s: %s" % (_STATUS.get(status, status), msg))
        self.status = status


class InstanceInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("nnz", C.c_uint32), ("graph_nnz", C.c_uint32), ("graph_nodes", C.c_uint32),
                ("nchi", C.c_uint32), ("offset", C.c_float), ("qubo", C.c_int32)]


class InstanceDesc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("nnz", C.c_uint32), ("row_ptr", _u32p), ("col", _u32p), ("J", _f32p),
                ("bias", _f32p), ("offset", C.c_float), ("nchi", C.c_uint32), ("suscept", _f64p),
                ("qubo", C.c_int32)]


class PtState(C.Structure):
    _fields_ = [("states", _i8p), ("energies", _f32p), ("num_acceptances", _u32p), ("round_trips", _u32p),
                ("tags", _u8p), ("diffusion_hist", _u32p)]


class PtParams(C.Structure):
    _fields_ = [("num_sweeps", C.c_uint32), ("warmup_fraction", C.c_double), ("num_betas", C.c_uint32),
                ("betas", _f32p), ("lo_beta_idx", C.c_uint32), ("icm", C.c_int32),
                ("num_replica_chains", C.c_uint32), ("threads", C.c_uint32), ("sample_enabled", C.c_int32),
                ("sample", C.c_uint32), ("sample_states_enabled", C.c_int32), ("sample_states", C.c_uint32),
                ("sample_limiting", C.c_uint32), ("seed", C.c_uint64), ("sweep_order", C.c_uint32)]


class PtResult(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("B", C.c_uint32), ("C", C.c_uint32), ("T", C.c_uint32), ("meas_init", C.c_uint32),
        ("nchi", C.c_uint32), ("npairs", C.c_uint32), ("n_meas", C.c_uint32),
        ("e", _f32p), ("q", _i32p), ("suscept", _f32p),
        ("n_samp_slots", C.c_uint32), ("n_samp_events", C.c_uint32), ("nbytes", C.c_uint32), ("samples", _u8p),
        ("n_gs", C.c_uint32), ("gs_words", C.c_uint32), ("gs_steps", _u32p), ("gs_energies", _f32p),
        ("gs_states", _u64p), ("acceptance_counts", _u32p), ("final_states", _i8p), ("final_energies", _f32p),
        ("final_ids", _u32p), ("num_acceptances", _u32p), ("round_trips", _u32p), ("tags", _u8p),
        ("diffusion_hist", _u32p), ("icm_moves", C.c_uint64), ("attempts", C.c_uint64), ("timing_us", C.c_double),
        ("device_ms", C.c_double),
    ]


class BatchOptions(C.Structure):
    _fields_ = [("waves_per_group", C.c_uint32), ("sweeps_per_launch", C.c_uint32), ("use_graph", C.c_int32),
                ("autotune", C.c_int32), ("task_tiling", C.c_uint32),
                ("icm_growth", C.c_uint32), ("teams", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("wall_ms", C.c_double), ("attempts", C.c_uint64),
                ("waves_per_group", C.c_uint32), ("sweeps_per_launch", C.c_uint32), ("launches", C.c_uint32),
                ("groups", C.c_uint32), ("phase_ms", C.c_double * 7), ("icm_passes", C.c_double),
                ("prep_ms", C.c_double), ("post_ms", C.c_double), ("groups_busy_ms", C.c_double),
                ("task_tiling", C.c_uint32), ("graph", C.c_int32), ("icm_growth", C.c_uint32),
                ("teams", C.c_uint32), ("work_groups", C.c_uint32)]


class SaParams(C.Structure):
    _fields_ = [("num_betas", C.c_uint32), ("betas", _f32p), ("num_replicas", C.c_uint32), ("threads", C.c_uint32),
                ("seed", C.c_uint64)]


class SaResult(C.Structure):
    _fields_ = [("n", C.c_uint32), ("R", C.c_uint32), ("min_energy", C.c_float), ("energies", _f32p),
                ("final_states", _i8p), ("attempts", C.c_uint64), ("timing_us", C.c_double),
                ("device_ms", C.c_double)]


class ServerOptions(C.Structure):
    _fields_ = [("pool_bytes", C.c_uint64), ("max_wave_slots", C.c_uint32), ("slice_ms", C.c_double),
                ("waves_per_group", C.c_uint32)]


class RequestTimes(C.Structure):
    _fields_ = [("queued_ms", C.c_double), ("total_ms", C.c_double), ("slices", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("cancelled", C.c_uint64), ("queued", C.c_uint32), ("resident", C.c_uint32),
                ("resident_waves", C.c_uint32), ("slices", C.c_uint64), ("device_busy_ms", C.c_double),
                ("attempts", C.c_uint64), ("pool_used", C.c_uint64), ("pool_bytes", C.c_uint64),
                ("slice_wall_ms", C.c_double), ("complete_ms", C.c_double)]


class Partition(C.Structure):
    _fields_ = [("num_partitions", C.c_uint32), ("offsets", _u32p), ("sites", _u32p)]


class GlaResult(C.Structure):
    _fields_ = [("instance_size", C.c_uint32), ("num_partitions", C.c_uint32), ("num_boundaries", C.c_uint32),
                ("partition_energies", _f32p), ("boundary_pairs", _u32p), ("boundary_energies", _f32p),
                ("bulk_energy", C.c_float), ("boundary_energy", C.c_float), ("total_energy", C.c_float),
                ("joined_state", _i8p), ("attempts", C.c_uint64), ("device_ms", C.c_double)]


_vp = C.c_void_p
_vpp = C.POINTER(C.c_void_p)
SIGNATURES = {
    "tamc_last_error": (C.c_char_p, []),
    "tamc_version": (C.c_int, []),
    "tamc_init": (C.c_int, []),
    "tamc_instance_from_file": (C.c_int, [C.c_char_p, C.c_int, _vpp]),
    "tamc_instance_from_lines": (C.c_int, [C.c_uint64, _u32p, _u32p, _f64p, C.c_int, _vpp]),
    "tamc_instance_create": (C.c_int, [C.POINTER(InstanceDesc), _vpp]),
    "tamc_instance_set_suscept": (C.c_int, [_vp, C.c_uint32, _f64p]),
    "tamc_instance_add_suscept_file": (C.c_int, [_vp, C.c_char_p]),
    "tamc_instance_get_info": (C.c_int, [_vp, C.POINTER(InstanceInfo)]),
    "tamc_instance_get_csr": (C.c_int, [_vp, _u32p, _u32p, _f32p, _f32p]),
    "tamc_instance_get_graph": (C.c_int, [_vp, _u32p, _u32p]),
    "tamc_instance_induced_subgraph": (C.c_int, [_vp, _u32p, C.c_uint32, _vpp]),
    "tamc_instance_destroy": (None, [_vp]),
    "tamc_energy": (C.c_int, [_vp, C.c_uint32, _i8p, _f32p]),
    "tamc_delta_energy": (C.c_int, [_vp, C.c_uint32, _i8p, _f32p]),
    "tamc_overlap": (C.c_int, [C.c_uint32, C.c_uint32, _i8p, _i8p, _i64p]),
    "tamc_pack_bytes": (C.c_int, [C.c_uint32, C.c_uint32, _i8p, _u8p]),
    "tamc_houdayer": (C.c_int, [_vp, C.c_uint32, _i8p, _i8p, _u32p, _i64p, _u8p]),
    "tamc_pt_resolve": (C.c_int, [C.c_int32, C.c_float, C.c_float, C.c_uint32, _f32p, C.c_uint32, C.c_float,
                                  C.c_int64, C.c_int32, _f32p, _u32p, _u32p]),
    "tamc_pt_run": (C.c_int, [_vp, C.POINTER(PtParams), _i8p, C.POINTER(C.POINTER(PtResult))]),
    "tamc_pt_run_lockstep": (C.c_int, [_vp, C.POINTER(PtParams), _i8p, C.c_uint32, C.POINTER(_u32p), _u64p,
                                       C.POINTER(C.POINTER(PtResult))]),
    "tamc_pt_run_resume": (C.c_int, [_vp, C.POINTER(PtParams), C.POINTER(PtState),
                                     C.POINTER(C.POINTER(PtResult))]),
    "tamc_pt_result_free": (None, [C.POINTER(PtResult)]),
    "tamc_pt_run_batch": (C.c_int, [C.c_uint32, _vpp, C.POINTER(PtParams), C.POINTER(BatchOptions),
                                    C.POINTER(C.POINTER(PtResult)), C.POINTER(BatchStats)]),
    "tamc_pt_run_plain": (C.c_int, [C.c_uint32, _vpp, C.POINTER(PtParams), C.POINTER(C.POINTER(PtResult)),
                                    C.POINTER(BatchStats)]),
    "tamc_pt_run_batch_init": (C.c_int, [C.c_uint32, _vpp, C.POINTER(PtParams), C.POINTER(_i8p),
                                         C.POINTER(BatchOptions), C.POINTER(C.POINTER(PtResult)),
                                         C.POINTER(BatchStats)]),
    "tamc_pt_opt_update": (C.c_int, [C.c_uint32, _f32p, _f32p, C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32,
                                     C.c_float, C.c_float, _f32p, _f32p, _f32p, C.POINTER(C.c_int)]),
    "tamc_server_create": (C.c_int, [C.POINTER(ServerOptions), _vpp]),
    "tamc_server_submit": (C.c_int, [_vp, _vp, C.POINTER(PtParams), _i8p, _u64p]),
    "tamc_server_poll": (C.c_int, [_vp, C.c_uint64, C.c_int32, C.POINTER(C.POINTER(PtResult)),
                                   C.POINTER(RequestTimes)]),
    "tamc_server_submit_sa": (C.c_int, [_vp, _vp, C.POINTER(SaParams), _u64p]),
    "tamc_server_poll_sa": (C.c_int, [_vp, C.c_uint64, C.c_int32, C.POINTER(C.POINTER(SaResult)),
                                      C.POINTER(RequestTimes)]),
    "tamc_server_cancel": (C.c_int, [_vp, C.c_uint64]),
    "tamc_server_get_stats": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "tamc_server_destroy": (None, [_vp]),
    "tamc_sa_run": (C.c_int, [_vp, C.POINTER(SaParams), C.POINTER(C.POINTER(SaResult))]),
    "tamc_sa_run_batch": (C.c_int, [C.c_uint32, _vpp, C.POINTER(SaParams), C.POINTER(C.POINTER(SaResult)),
                                    C.POINTER(BatchStats)]),
    "tamc_sa_result_free": (None, [C.POINTER(SaResult)]),
    "tamc_sa_run_noalloc": (C.c_int, [_vp, C.POINTER(SaParams), C.c_void_p, C.c_void_p, C.c_void_p]),
    "tamc_write_thermal_samples": (C.c_int, [C.POINTER(PtResult), _f32p, C.c_uint8, C.c_char_p]),
    "tamc_write_anneal_state": (C.c_int, [C.POINTER(SaResult), C.c_char_p]),
    "tamc_write_thermal_samples_pickle": (C.c_int, [C.POINTER(PtResult), _f32p, C.c_uint8, C.c_char_p]),
    "tamc_write_anneal_state_pickle": (C.c_int, [C.POINTER(SaResult), C.c_char_p]),
    "tamc_partition_from_file": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.POINTER(Partition))]),
    "tamc_partition_create": (C.c_int, [_vp, C.c_uint32, _u32p, _u32p, C.POINTER(C.POINTER(Partition))]),
    "tamc_partition_free": (None, [C.POINTER(Partition)]),
    "tamc_percolate": (C.c_int, [_vp, C.c_double, C.c_uint32, C.c_uint64, _u32p, _u8p]),
    "tamc_gla_run": (C.c_int, [_vp, C.POINTER(Partition), C.POINTER(PtParams), C.POINTER(C.POINTER(GlaResult))]),
    "tamc_gla_result_free": (None, [C.POINTER(GlaResult)]),
    "tamc_write_gla_results": (C.c_int, [C.POINTER(GlaResult), C.c_char_p]),
}

_lib = None


def lib():
    """Load libtamc_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TamcError(TAMC_E_DEVICE, "libtamc_gpu.so not built: run python -m tamc_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != TAMC_OK:
        raise TamcError(status, lib().tamc_last_error().decode(errors="replace"))


def ptr(a, t):
    return a.ctypes.data_as(t)
