"""tamc-pt-opt's temperature-set update (src/bin/tamc-pt-opt.rs:57-167, tamc-core/src/util.rs:6-138)
in the library against the float32 restatement in oracle/pt_opt.py, and the restatement's helpers
against closed forms."""
import ctypes as C

import numpy as np
import pytest

from oracle import pt_opt as OP
from tamc_amd import _lib

f32 = np.float32


def _geom(bmin, bmax, nt):
    return [f32(np.exp(np.log(bmin) + k * (np.log(bmax) - np.log(bmin)) / (nt - 1))) for k in range(nt)]


def _inputs(rng, nt, count, sweeps=2000, chains=2):
    beta = _geom(0.2, 5.0, nt)
    hists, rts = [], []
    for _ in range(count):
        frac = 1.0 / (1.0 + np.exp(-(np.array(beta, np.float64) - 1.2) / 0.4))
        frac = np.clip(frac + rng.normal(0, 0.02, nt), 0, 1)
        tot = rng.integers(500, 3000, nt)
        h1 = np.round(tot * frac).astype(np.uint32)
        hists.append(np.stack([tot.astype(np.uint32) - h1, h1], axis=1))
        rts.append(int(rng.integers(0, 60)))
    return beta, hists, rts


def _lib_update(beta, mom, hists, rts, sweeps, chains, alpha, m):
    nt, count = len(beta), len(hists)
    b = np.array(beta, np.float32)
    mo = np.array(mom, np.float32)
    h = np.ascontiguousarray(np.stack(hists), np.uint32)
    r = np.array(rts, np.uint32)
    out = np.zeros(nt, np.float32)
    taus = np.zeros(count, np.float32)
    err = C.c_float()
    upd = C.c_int()
    _lib.check(_lib.lib().tamc_pt_opt_update(nt, _lib.ptr(b, _lib._f32p), _lib.ptr(mo, _lib._f32p), count,
                                             _lib.ptr(h, _lib._u32p), _lib.ptr(r, _lib._u32p), sweeps, chains,
                                             alpha, m, _lib.ptr(out, _lib._f32p), C.byref(err),
                                             _lib.ptr(taus, _lib._f32p), C.byref(upd)))
    return out, mo, f32(err.value), taus, bool(upd.value)


@pytest.mark.parametrize("nt,count,seed", [(16, 4, 1), (32, 8, 2), (9, 1, 3)])
def test_update_matches_restatement(nt, count, seed):
    rng = np.random.default_rng(seed)
    beta, hists, rts = _inputs(rng, nt, count)
    rts[0] = max(rts[0], 1)  # at least one round trip so that tau and eta are not all zero
    mom = list(beta)
    for it in range(3):  # the momentum carries over between iterations
        want = OP.update(beta, mom, hists, rts, 2000, 2, 0.2, 0.85)
        got = _lib_update(beta, mom, hists, rts, 2000, 2, 0.2, 0.85)
        assert want is not None and got[4]
        assert np.array_equal(got[0], np.array(want[0], np.float32))
        assert np.array_equal(got[1], np.array(want[1], np.float32))
        assert got[2] == want[2]
        assert np.array_equal(got[3], np.array(want[3], np.float32))
        beta, mom = list(got[0]), list(got[1])
        # endpoints stay at the ends; the set stays increasing
        assert np.all(np.diff(got[0]) > 0)


def test_update_without_round_trips_keeps_the_set():
    rng = np.random.default_rng(4)
    beta, hists, rts = _inputs(rng, 12, 3)
    rts = [0, 0, 0]
    assert OP.update(beta, list(beta), hists, rts, 2000, 2, 0.2, 0.85) is None
    out, mo, err, taus, upd = _lib_update(beta, list(beta), hists, rts, 2000, 2, 0.2, 0.85)
    assert not upd and np.array_equal(out, np.array(beta, np.float32)) and np.all(taus == 0)
    assert np.array_equal(mo, np.array(beta, np.float32))


def test_restatement_helpers_closed_forms():
    x = [f32(v) for v in (0.0, 1.0, 3.0, 6.0)]
    # trapezoid weights and a second-order exact derivative of a quadratic
    assert OP.stepwise_weights(x) == [f32(0.5), f32(1.5), f32(2.5), f32(1.5)]
    y = [f32(v * v) for v in x]
    d = OP.finite_differences(x, y)
    assert d[1] == f32(2.0) and d[2] == f32(6.0)  # exact at interior points for a quadratic
    # ndarray's unrolled sum equals the exact sum for small integers
    assert OP.unrolled_sum([f32(k) for k in range(21)]) == f32(210)
    # monotonic divisions of the identity are the equal parts; of x^2 their square roots
    xd = OP.monotonic_divisions(lambda t: f32(t), 4)
    assert np.allclose(xd, [0, 0.25, 0.5, 0.75, 1], atol=2e-6)
    xd = OP.monotonic_divisions(lambda t: f32(t * t), 4)
    assert np.allclose(xd, np.sqrt([0, 0.25, 0.5, 0.75, 1]), atol=2e-6)


def test_pt_opt_command_line_arguments():
    import os
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tamc_amd", "lib", "tamc-pt-opt")
    r = subprocess.run([cli, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Optimize the PT temperature set" in r.stdout and "--tolerance" in r.stdout
    r = subprocess.run([cli, "--bogus", "x.yml"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--bogus" in r.stderr
    r = subprocess.run([cli], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "<params>" in r.stderr
