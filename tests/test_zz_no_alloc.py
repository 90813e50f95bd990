"""The allocation-free SA entry point (tamc_sa_run_noalloc) against tamc_sa_run on the same
inputs, through the C ABI: two calls in a row, the second on a smaller input, bit-exact; and the
size limit of its scratch."""
import ctypes as C

import numpy as np
import pytest

from tamc_amd import _lib, api
from tests.helpers import make_pair

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libtamc_gpu.so is linked against (the test allocates the output buffers
    itself: the entry point must not)."""
    _lib.lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


def _noalloc(inst, params, R):
    hip = _hip()
    betas = params.betas()
    P = _lib.SaParams(len(betas), betas.ctypes.data_as(_lib._f32p), params.num_replicas, params.threads, params.seed)
    de, ds = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(de), max(R, 1) * 4) == 0 and hip.hipMalloc(C.byref(ds), max(R * inst.n, 1)) == 0
    try:
        assert hip.hipMemset(de, 0xFF, max(R, 1) * 4) == 0 and hip.hipMemset(ds, 0, max(R * inst.n, 1)) == 0
        st = _lib.lib().tamc_sa_run_noalloc(inst._h, C.byref(P), de, ds, None)
        assert hip.hipDeviceSynchronize() == 0
        _lib.check(st)
        e = np.zeros(R, np.float32)
        s = np.zeros((R, inst.n), np.int8)
        assert hip.hipMemcpy(e.ctypes.data, de, R * 4, 2) == 0  # device to host
        assert hip.hipMemcpy(s.ctypes.data, ds, R * inst.n, 2) == 0
        return e, s
    finally:
        hip.hipFree(de)
        hip.hipFree(ds)


def test_noalloc_sa_equals_sa_run_twice(gpu):
    cases = [("qubo", dict(n=200, density=0.1), 300, 3, 60, 5), ("2d_gauss", dict(L=6), 100, 1, 30, 9)]
    for kind, kw, R, threads, K, seed in cases:  # the second call is smaller in n, R and K
        _, gi, _ = make_pair(kind, seed=seed, **kw)
        betas = np.linspace(0.1, 3.0, K).astype(np.float32)
        p = api.SaParams(beta=api.BetaOptions.Arr(betas), num_replicas=R, threads=threads, seed=seed)
        ref = api.SaRunner(gi, p).run()  # also uploads the instance
        e, s = _noalloc(gi, p, ref["R"])
        assert np.array_equal(e, ref["energies"]), kind
        assert np.array_equal(s, ref["final_states"]), kind


def test_noalloc_sa_rejects_what_it_cannot_hold(gpu):
    _, gi, _ = make_pair("2d_pm", seed=1, L=4)
    api.SaRunner(gi, api.SaParams(beta=api.BetaOptions.Arr([0.5, 1.0]), num_replicas=64, seed=1)).run()
    long = np.linspace(0.1, 1.0, 131073).astype(np.float32)
    p = api.SaParams(beta=api.BetaOptions.Arr(long), num_replicas=64, seed=1)
    with pytest.raises(_lib.TamcError) as e:
        _noalloc(gi, p, 64)
    assert e.value.status == _lib.TAMC_E_INVALID_ARG
    _, fresh, _ = make_pair("2d_pm", seed=2, L=4)  # never run: not on the device yet
    with pytest.raises(_lib.TamcError) as e:
        _noalloc(fresh, api.SaParams(beta=api.BetaOptions.Arr([0.5]), num_replicas=64, seed=1), 64)
    assert e.value.status == _lib.TAMC_E_INVALID_ARG
    _, big, _ = make_pair("rr3", seed=3, n=21000)  # its planes would need global memory: refused
    with pytest.raises(_lib.TamcError) as e:
        _noalloc(big, api.SaParams(beta=api.BetaOptions.Arr([0.5]), num_replicas=64, seed=1), 64)
    assert e.value.status == _lib.TAMC_E_UNSUPPORTED
