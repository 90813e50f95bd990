#!/usr/bin/env python3
"""Benchmark: spin-flip attempts/s of the batched persistent PT-ICM kernel on one MI355X.

Workload (the batched PT-ICM configuration): 256 random 3-regular graphs, N ~ U{200..4000} even,
J ~ N(0,1) and h ~ N(0, 0.1) as float32; 32 geometric betas 0.2-5.0, 2 replica chains, 4000 sweeps,
ICM on slots with beta >= 1.0, measurements every 32 sweeps after a 0.5 warm-up.  A "step" is one
complete batched solve of all 256 instances (T*N*B*C Metropolis attempts summed over instances).
Instance data is uploaded (and cached on the device) during warm-up, so the timed steps start with
inputs resident in HBM.

Reported beside it:
  * the plain path (one instance per kernel launch, run one after another) on a sample of instances;
  * the CPU restatement of the reference on the host cores (cpu_baseline: all the cores this process
    may run on, and one core);
  * the other configurations: the 2D L=16 and 3D L=12 single instances on the plain path, SA on 64
    QUBOs over the full 4096-sweep schedule, each with one-core and all-core CPU figures;
  * serving: the Poisson request stream at its configured rate (200/s) and at a rate below the
    server's capacity, p50/p99 time-to-result beside throughput;
  * for the dominant kernels, the VALU-issue roofline (wave-instructions per attempt from the
    committed rocprofv3 counter run, times the live attempt rate, against 1 wave-instruction per
    clock per compute unit) and how busy the mean compute unit is.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# aggregate LDS bandwidth assumed for the LDS-level fraction: 256 CUs x 128 B/clk x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4
CUS, CLOCK_HZ = 256, 2.4e9
# VALU issue: one wave64 instruction per clock per compute unit (4 SIMDs, 4 clocks per wave64 op)
VALU_PEAK = CUS * CLOCK_HZ
# kernel phases timed per group (tamc_batch_stats.phase_ms order)
PHASES = ["icm_overlap", "icm_seed", "icm_growth", "icm_flip", "exchange", "sweep", "measure"]
BYTES_PER_ATTEMPT_D3 = 9 * 3 + 9  # 4 B col + 4 B J + 1 B spin per edge; 4 B row_ptr + 4 B bias + 1 B spin per site


def host_cores():
    """The cores this process may run on (its affinity set), which is what the CPU legs use."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _profile(name):
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def measured_traffic():
    """HBM bytes per launch of pt_sweep_kernel on this workload, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs recorded beside the kernel-trace summary (scaled by the copy calibration)."""
    t = _profile("pt_sweep_kernel_rr3x256_traffic.json")
    return t.get("traffic_bytes_per_launch") if t else None


def valu_roofline(kernel, attempts_per_s):
    """VALU-issue roofline from the committed counter run: SQ_INSTS_VALU per attempt of `kernel` on
    its workload, times the live attempt rate, against VALU_PEAK."""
    c = _profile("kernel_counters.json")
    if not c or kernel not in c:
        return None
    per = c[kernel]["valu_per_attempt"]
    achieved = per * attempts_per_s
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK, "unit": "wave-instructions/s",
            "frac": achieved / VALU_PEAK, "valu_per_attempt": per, "counter_workload": c[kernel]["workload"]}


def build_workload(count, seed, T):
    from tamc_amd import api, workloads
    insts, lines = [], []
    for k, (i, j, K, n) in enumerate(workloads.config3_instances(count=count, seed=seed)):
        insts.append(api.BqmIsingInstance.from_lines(i, j, K))
        lines.append((i, j, K))
    params = [api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 32),
                              lo_beta=1.0, icm=True, num_replica_chains=2, threads=1, sample=32, sample_states=32,
                              sample_limiting=2, seed=1000 + k) for k in range(count)]
    return insts, lines, params


def cpu_baseline(lines, params, sweeps, threads):
    """The CPU restatement (reference algorithm: row gather per attempt, Xoshiro256++, serial
    pt_loop) on a bounded sample: every instance for `sweeps` sweeps, instances spread over threads."""
    import oracle as O
    betas = O.beta_geometric(0.2, 5.0, 32)
    lo = int(np.argmax(betas >= 1.0))
    insts = [O.Instance.from_lines(i, j, K) for (i, j, K) in lines]

    def one(k):
        r = O.pt_run(insts[k], betas, sweeps, num_chains=2, warmup_fraction=0.5, lo_beta_idx=lo, icm=True,
                     sample=32, sample_states=32, sample_limiting=2, rng_mode=O.RNG_XOSHIRO, seed=k)
        return r["attempts"]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        attempts = sum(ex.map(one, range(len(insts))))
    dt = time.perf_counter() - t0
    return attempts / dt, attempts, dt


def _cpu_parallel(fn, copies, threads):
    """Run fn(k) for k < copies on `threads` host threads; (attempts/s, attempts, seconds)."""
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        att = sum(ex.map(fn, range(copies)))
    dt = time.perf_counter() - t0
    return att / dt, att, dt


def serving_leg(reqs, insts, rate, sweeps, slice_ms):
    """Serving workload: the seeded stream of mixed PT-ICM requests (2D L 8..128, 3D L 4..24, random
    3-regular N 64..16384; +-J or N(0,1); 32 betas 0.2-5.0, 2 chains, `sweeps` sweeps) fed to the
    continuous-batching server at arrival times scaled to `rate` requests/s.  Latency = submit ->
    result on the host."""
    from tamc_amd import api
    base_rate = len(reqs) / reqs[-1][0]
    scale = 200.0 / rate  # the stream was drawn at 200/s; a Poisson process rescales in time

    def params(k):
        return api.PtIcmParams(num_sweeps=sweeps, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 32),
                               lo_beta=1.0, icm=True, num_replica_chains=2, sample=32, sample_states=32,
                               sample_limiting=2, seed=5000 + k)

    with api.Server(slice_ms=slice_ms) as srv:
        srv.wait(srv.submit(insts[0], params(0)))  # warm-up request
        tickets = []
        t0 = time.perf_counter()
        for k, (arrival, _, _, _) in enumerate(reqs):
            delay = arrival * scale - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(insts[k], params(k)))
        submitted_s = time.perf_counter() - t0
        lat, queued, attempts, info = [], [], 0, []
        for k, t in enumerate(tickets):
            res, tm = srv.wait(t)
            lat.append(tm["total_ms"])
            queued.append(tm["queued_ms"])
            attempts += res["attempts"]
            info.append({"request": k, "kind": reqs[k][1], "size": reqs[k][2], "arrival_s": round(reqs[k][0] * scale, 4),
                         "total_ms": tm["total_ms"], "queued_ms": tm["queued_ms"], "slices": tm["slices"]})
        makespan = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array(lat)
    return {
        "requests": len(reqs), "arrival_rate_per_s": rate, "realised_rate_per_s": base_rate / scale,
        "sweeps": sweeps, "slice_ms": slice_ms,
        "p50_ms": float(np.percentile(lat, 50)), "p90_ms": float(np.percentile(lat, 90)),
        "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
        "queued_p50_ms": float(np.percentile(queued, 50)), "queued_p99_ms": float(np.percentile(queued, 99)),
        "throughput_attempts_per_s": attempts / makespan, "makespan_s": makespan, "arrivals_s": submitted_s,
        "offered_attempts_per_s": attempts / (reqs[-1][0] * scale), "slices": st["slices"],
        "device_busy_s": st["device_busy_ms"] / 1e3, "slice_wall_s": st["slice_wall_ms"] / 1e3,
        "complete_s": st["complete_ms"] / 1e3,
        # the slowest requests: shape, arrival, wait for admission and slices taken
        "slowest": sorted(info, key=lambda x: -x["total_ms"])[:4],
    }


def other_configs(with_cpu, cores, sa_sweeps):
    """The reference's other configurations (see pt_plain_cases and sa_case) and the largest
    serving shapes alone on the device."""
    out = pt_plain_cases(with_cpu, cores)
    out["sa_qubo"] = sa_case(with_cpu, cores, sa_sweeps)
    out["largest_serving_shapes"] = largest_serving_shapes()
    return out


def pt_plain_cases(with_cpu, cores):
    """The reference's other configurations, each next to the CPU restatement (the reference's
    serial loop with Xoshiro256++) on a bounded sample, on one core and on all `cores` (that many
    independent copies of the sample at once: the throughput the host's cores give on the shape):
      plain_2d: 2D L=16 periodic +-J, 32 betas 0.2-5.0, 2 chains, 2000 sweeps (one instance, plain path)
      plain_3d: 3D L=12 periodic N(0,1), 48 betas 0.1-5.0, 4 chains, 10000 sweeps (one instance)
      sa_qubo:  64 QUBOs N=2048, 5% density, Q ~ U{-100..100}, 1024 restarts each, linear beta
                0.1 -> 10 over 4096 sweeps (the full schedule; `sa_sweeps` evenly spaced entries of it
                when smaller)."""
    from tamc_amd import api, workloads as W
    out = {}

    def pt_case(name, lines, B, bmin, bmax, C, T, cpu_sweeps):
        i, j, K, n = lines
        inst = api.BqmIsingInstance.from_lines(i, j, K)
        p = api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(bmin, bmax, B),
                            lo_beta=1.0, icm=True, num_replica_chains=C, sample=32, sample_states=32,
                            sample_limiting=2, seed=1)
        api.PtIcmRunner(inst, p).run()  # warm (instance upload)
        t0 = time.perf_counter()
        _, st = api.run_pt_batch([inst], [p], plain=True)
        dt = time.perf_counter() - t0
        rec = {"value": st["attempts"] / dt, "unit": "spin-flip attempts/s", "attempts": int(st["attempts"]),
               "ms": dt * 1e3, "kernel_ms": st["device_ms"], "n": n, "work_groups": st["work_groups"],
               "teams": st["teams"],
               "mean_cu_busy": st["groups_busy_ms"] / (CUS * st["device_ms"]) if st["device_ms"] else None}
        if with_cpu:
            import oracle as O
            betas, lo = p.resolve()

            def one(k):
                oi = O.Instance.from_lines(i, j, K)
                return O.pt_run(oi, betas, cpu_sweeps, num_chains=C, warmup_fraction=0.5, lo_beta_idx=lo, icm=True,
                                sample=32, sample_states=32, sample_limiting=2, rng_mode=O.RNG_XOSHIRO,
                                seed=1 + k)["attempts"]
            r1, a1, d1 = _cpu_parallel(one, 1, 1)
            ra, aa, da = _cpu_parallel(one, cores, cores)
            rec["cpu_1thread"] = {"value": r1, "cores": 1, "sample": "%d of %d sweeps (%.1f s)" % (cpu_sweeps, T, d1)}
            rec["cpu_allcore"] = {"value": ra, "cores": cores,
                                  "sample": "%d independent copies of %d of %d sweeps at once (%.1f s)"
                                            % (cores, cpu_sweeps, T, da)}
            rec["vs_cpu_1thread"] = rec["value"] / r1
            rec["vs_cpu_allcore"] = rec["value"] / ra
        out[name] = rec

    pt_case("plain_2d", W.lattice_2d(16, np.random.default_rng(101), pm_j=True), 32, 0.2, 5.0, 2, 2000, 2000)
    pt_case("plain_3d", W.lattice_3d(12, np.random.default_rng(102), pm_j=False), 48, 0.1, 5.0, 4, 10000, 100)
    return out


def sa_case(with_cpu, cores, sa_sweeps):
    """SA on 64 QUBOs (N=2048, 5 %), 1024 restarts each, the linear 0.1 -> 10 schedule of 4096
    sweeps (or `sa_sweeps` evenly spaced entries of it), with one-core and all-core CPU figures."""
    from tamc_amd import api, workloads as W
    insts, lines = [], []
    for k in range(64):
        q = W.qubo_sparse(2048, 0.05, np.random.default_rng(104 + k), qmax=100)
        lines.append(q)
        insts.append(api.BqmIsingInstance.from_lines(q[0], q[1], q[2], qubo=True))
    betas = (0.1 + 9.9 * np.arange(4096) / 4095).astype(np.float32)
    sched = betas if sa_sweeps >= 4096 else betas[np.linspace(0, 4095, sa_sweeps).astype(int)]
    ps = [api.SaParams(beta=api.BetaOptions.Arr(sched), num_replicas=1024, threads=1, seed=k) for k in range(64)]
    api.run_sa_batch(insts[:1], [api.SaParams(beta=api.BetaOptions.Arr(betas[:4]), num_replicas=64, seed=0)])  # warm
    t0 = time.perf_counter()
    _, st = api.run_sa_batch(insts, ps)
    dt = time.perf_counter() - t0
    kernel_rate = st["attempts"] / (st["device_ms"] / 1e3)
    rec = {"value": st["attempts"] / dt, "unit": "spin-flip attempts/s", "attempts": int(st["attempts"]),
           "ms": dt * 1e3, "kernel_ms": st["device_ms"], "kernel_attempts_per_s": kernel_rate,
           "sample": "all 64 instances x 1024 restarts, %d of the 4096-sweep schedule" % len(sched),
           # 9 d + 9 algorithmic bytes per attempt at d ~ 103.4 (QUBO self entry included): 939 B
           "roofline_hbm": {"achieved": kernel_rate * 939 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": kernel_rate * 939 / 1e9 / HBM_PEAK_GBS},
           "roofline_valu": valu_roofline("sa_kernel", kernel_rate)}
    if with_cpu:
        import oracle as O
        S = 64

        def one(k):
            q = lines[k % 64]
            oi = O.Instance.from_lines(q[0], q[1], q[2], qubo=True)
            return O.sa_run(oi, betas[:S], 64, threads=1, rng_mode=O.RNG_XOSHIRO, seed=1 + k)["attempts"]
        r1, a1, d1 = _cpu_parallel(one, 1, 1)
        ra, aa, da = _cpu_parallel(one, cores, cores)
        rec["cpu_1thread"] = {"value": r1, "cores": 1, "sample": "1 instance x 64 restarts x %d sweeps (%.1f s)" % (S, d1)}
        rec["cpu_allcore"] = {"value": ra, "cores": cores,
                              "sample": "%d instances x 64 restarts x %d sweeps at once (%.1f s)" % (cores, S, da)}
        rec["vs_cpu_1thread"] = rec["value"] / r1
        rec["vs_cpu_allcore"] = rec["value"] / ra
    return rec


def largest_serving_shapes():
    """The largest request shapes of the serving stream, each alone on the device."""
    from tamc_amd import api, workloads as W
    # the largest request shapes of the serving stream, one instance alone on the device (serving
    # parameters: 32 betas 0.2-5.0, 2 chains, ICM), timed over 100 sweeps: the solve latency floor
    # of the stream's tail, with the kernel's per-phase split
    shapes = {"2d_L128_pmJ": W.lattice_2d(128, np.random.default_rng(8), pm_j=True),
              "3d_L24_gauss": W.lattice_3d(24, np.random.default_rng(7)),
              "rr3_N16384_gauss": W.random_regular3(16384, np.random.default_rng(5), gaussian=True)}
    largest = {}
    for name, (i, j, K, n) in shapes.items():
        inst = api.BqmIsingInstance.from_lines(i, j, K)
        T = 100
        p = api.PtIcmParams(num_sweeps=T, warmup_fraction=0.5, beta=api.BetaOptions.Geometric(0.2, 5.0, 32),
                            lo_beta=1.0, icm=True, num_replica_chains=2, sample=32, sample_states=32,
                            sample_limiting=2, seed=1)
        api.run_pt_batch([inst], [p], waves_per_group=16)  # warm (instance upload)
        _, st = api.run_pt_batch([inst], [p], waves_per_group=16)
        largest[name] = {"n": n, "us_per_sweep": st["device_ms"] * 1e3 / T,
                         "phases_us_per_sweep": {k: v * 1e3 / T for k, v in zip(PHASES, st["phase_ms"])},
                         "icm_passes_per_sweep": st["icm_passes"] / T}
    return largest


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--sweeps", type=int, default=4000)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--cpu-sweeps", type=int, default=200)
    ap.add_argument("--cpu-instances", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--sweeps-per-launch", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true", help="launch without hipGraph capture")
    ap.add_argument("--tiling", type=int, default=0, help="task tiling (1 dealt in turn, 2 runs; 0 = autotuned)")
    ap.add_argument("--teams", type=int, default=0, help="teams of work groups (1 never, K forced; 0 = autotuned)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-other", action="store_true", help="skip the other configurations")
    ap.add_argument("--sa-sweeps", type=int, default=4096, help="SA schedule entries timed (4096 = all)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-low-rate", type=float, default=50.0,
                    help="second serving rate, below the server's capacity (0 = skip)")
    ap.add_argument("--serve-sweeps", type=int, default=2000)
    ap.add_argument("--serve-slice-ms", type=float, default=4.0)
    ap.add_argument("--only-largest", action="store_true", help="run only the largest instance (profiling)")
    ap.add_argument("--no-icm", action="store_true", help="switch ICM off (profiling the sweep alone)")
    ap.add_argument("--only-sa", action="store_true", help="run only the SA configuration (profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (float32 elementwise), the reference point "
                         "that scales rocprofv3 FETCH_SIZE/WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": args.copy_calibration, "read_bytes": 4 * n, "write_bytes": 4 * n}))
        return 0
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (independent instances would shard across GPUs with "
              "no communication); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2

    from tamc_amd import api
    from tamc_amd._lib import check, lib
    check(lib().tamc_init())
    cores = host_cores()
    if args.only_sa:
        print(json.dumps({"sa_qubo": sa_case(False, cores, args.sa_sweeps)}))
        return 0
    insts, lines, params = build_workload(args.instances, 103, args.sweeps)
    if args.only_largest:
        k = int(np.argmax([x.n for x in insts]))
        insts, lines, params = [insts[k]], [lines[k]], [params[k]]
    if args.no_icm:
        for p in params:
            p.icm = False
    sum_n = sum(x.n for x in insts)
    attempts_per_step = sum(p.num_sweeps * x.n * 32 * 2 for p, x in zip(params, insts))

    def step():
        return api.run_pt_batch(insts, params, waves_per_group=args.waves, sweeps_per_launch=args.sweeps_per_launch,
                                use_graph=not args.no_graph, autotune=True, task_tiling=args.tiling,
                                teams=args.teams)

    for _ in range(args.warmup):
        step()
    dev_ms, prep_ms, post_ms, busy = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, st = step()
        dev_ms.append(st["device_ms"])
        prep_ms.append(st["prep_ms"])
        post_ms.append(st["post_ms"])
        busy.append(st["groups_busy_ms"] / (CUS * st["device_ms"]))
        assert st["attempts"] == attempts_per_step
    wall = time.perf_counter() - t0
    value = attempts_per_step * args.steps / wall
    kernel_rate = attempts_per_step / (np.mean(dev_ms) / 1e3)
    ms_per_step = wall / args.steps * 1e3

    # plain path: one instance per launch, sequentially, on a sample of instances
    plain_rate = None
    k = min(args.plain_sample, len(insts))
    if not args.no_plain and k > 0:
        _, pst = api.run_pt_batch(insts[:1], params[:1], plain=True)  # warm
        t1 = time.perf_counter()
        _, pst = api.run_pt_batch(insts[:k], params[:k], plain=True)
        plain_wall = time.perf_counter() - t1
        plain_rate = pst["attempts"] / plain_wall

    achieved = kernel_rate * BYTES_PER_ATTEMPT_D3 / 1e9
    default_workload = (args.instances == 256 and args.sweeps == 4000 and not args.only_largest
                        and not args.no_icm)
    out = {
        "metric": "spin-flip attempts/s per GPU (batched PT-ICM vs plain) + p50/p99 solve latency",
        "value": value,
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched PT-ICM: %d random 3-regular graphs N~U{200..4000} (sum N=%d), J~N(0,1), "
                        "h~N(0,0.1), 32 geometric betas 0.2-5.0, 2 chains, %d sweeps, ICM beta>=1.0"
                        % (len(insts), sum_n, args.sweeps),
            "instances": len(insts), "sum_n": sum_n, "betas": 32, "chains": 2, "sweeps": args.sweeps,
            "waves_per_group": st["waves_per_group"], "sweeps_per_launch": st["sweeps_per_launch"],
            "launches": st["launches"], "task_tiling": {1: "dealt in turn", 2: "contiguous runs"}[st["task_tiling"]],
            "hipgraph": st["graph"],
            "icm_growth": {0: "per-instance default", 1: "strided sites", 2: "compact tiles"}[st["icm_growth"]],
            "teams": st["teams"], "work_groups": st["work_groups"],
        },
        "kernel_attempts_per_s": kernel_rate,
        # where a step's time outside the kernel goes: library preparation (planning, allocation,
        # uploads, state init) and post-processing (downloads, assembly); the rest is Python
        "host_ms_per_step": {"prep": float(np.mean(prep_ms)), "post": float(np.mean(post_ms)),
                             "other": ms_per_step - float(np.mean(dev_ms)) - float(np.mean(prep_ms))
                             - float(np.mean(post_ms))},
        # the launch lasts as long as its slowest group (one instance per CU): that group's phases
        "slowest_group_us_per_sweep": {k2: v * 1e3 / args.sweeps for k2, v in zip(PHASES, st["phase_ms"])},
        "slowest_group_icm_passes_per_sweep": st["icm_passes"] / args.sweeps,
        "plain_gpu": {"value": plain_rate, "unit": "spin-flip attempts/s", "instances": k,
                      "note": "one instance per kernel launch, launched one after another"},
        "batched_vs_plain": (value / plain_rate) if plain_rate else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic() if default_workload else None,
                     "bytes_per_attempt": BYTES_PER_ATTEMPT_D3,
                     "algorithmic_bytes_per_launch": attempts_per_step * BYTES_PER_ATTEMPT_D3,
                     "kernel_ms": float(np.mean(dev_ms)),
                     # the state lives in LDS for the whole launch: the same bytes against LDS
                     "lds_level": {"peak": LDS_PEAK_GBS, "unit": "GB/s", "frac": achieved / LDS_PEAK_GBS},
                     # what does bound it: VALU issue, and the share of the launch the mean CU is busy
                     "valu": valu_roofline("pt_sweep_kernel", kernel_rate) if default_workload else None,
                     "mean_cu_busy": float(np.mean(busy))},
    }
    if not args.no_other:
        out["other_configs"] = other_configs(not args.no_cpu, cores, args.sa_sweeps)
    if not args.no_serving:
        from tamc_amd import workloads
        reqs = workloads.serving_requests(args.serve_requests, 200.0, seed=105)
        sinsts = [api.BqmIsingInstance.from_lines(i, j, K) for (_, _, _, (i, j, K, n)) in reqs]
        out["serving"] = serving_leg(reqs, sinsts, args.serve_rate, args.serve_sweeps, args.serve_slice_ms)
        if args.serve_low_rate > 0:
            out["serving_below_capacity"] = serving_leg(reqs, sinsts, args.serve_low_rate, args.serve_sweeps,
                                                        args.serve_slice_ms)
    if not args.no_cpu:
        rate, att, dt = cpu_baseline(lines[:args.cpu_instances], params[:args.cpu_instances], args.cpu_sweeps,
                                     cores)
        out["cpu_baseline"] = {"value": rate, "unit": "spin-flip attempts/s", "cores": cores, "kind": "port",
                               "sample": "%d instances x %d sweeps of the same workload (%.3g attempts, %.1f s)"
                                         % (min(args.cpu_instances, len(lines)), args.cpu_sweeps, att, dt)}
        out["batched_vs_cpu"] = value / rate
        rate1, att1, dt1 = cpu_baseline(lines[:16], params[:16], 50, 1)
        out["cpu_baseline_1thread"] = {"value": rate1, "unit": "spin-flip attempts/s", "cores": 1,
                                       "sample": "16 instances x 50 sweeps (%.3g attempts, %.1f s)" % (att1, dt1)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
