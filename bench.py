#!/usr/bin/env python
"""Benchmark of the per-step hot path on one MI355X.

Workload (default): SyMBA, central body of 1 M_sun plus 32,768 fully
interacting massive bodies (masses log-uniform 1e-7..1e-3 M_sun, radii from
3 g/cm^3, a U(0.5,5) AU, e U(0,0.05), i U(0,2 deg)) and 1,048,576 test
particles (a U(0.3,10) AU, e U(0,0.2), i U(0,10 deg)); dt = 0.005 yr, FP64,
snapshots every 100 steps, all state resident in HBM.

`value` is FP64 pair interactions per second of whole-step wall time:
2 evaluations x (N(N-1)/2 + npl*ntp) pair interactions per step, divided by
the measured time of complete steps (encounter detection, recursion, drifts
and frame changes included).  The steps run with the fast pair factor
(SWG_PAIR_FAST: within a few ulp per pair, accelerations within 1e-13 of the
exact sums); the exact-factor kernels are timed on the same state beside it.
The roofline object covers the dominant kernel (pl-tp accelerations, 17
flops per pair) with its launch durations measured by HIP events on its own
stream during the timed steps.  The CPU baseline times the CPU restatement's
OpenMP pair kernels on a bounded sample of the same state (the unit of
`value`), and whole runs of the two small workloads (WHM 8 planets + 1000
particles, 10^4 steps; the 2000-body SyMBA disk, 10^3 steps) beside the same
runs on the GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("FP64 pair-interactions/s + steps/s at N=32k; ensemble system-steps/s, p50/p99")
PEAK_FP64_TFLOPS = 78.6        # MI355X FP64 vector
FLOPS_PER_PLTP_PAIR = 17       # swiftest_kick.f90:404-408, :464-467
FLOPS_PER_PLPL_PAIR = 19       # swiftest_kick.f90:227-236, per ordered pair


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/<kernel>_c3_hbm_traffic.md),
    or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_c3_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(sysic, target_s):
    """Pair-interaction rate of the CPU restatement's OpenMP kernels (the
    reference's loop decomposition: OpenMP over particles / rows, j in order)
    on a bounded sample of the benchmark state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O  # the test oracle doubles as the CPU baseline
    cores = cpu_threads()
    npl = sysic.npl
    rpl, Gm = sysic.pl_rh, sysic.Gmass
    # calibrate on a small slice, then size the timed slice to ~target_s
    ncal = 2048
    t0 = time.perf_counter()
    O.accel_pltp(sysic.tp_rh[:ncal], rpl, Gm, omp=True)
    tcal = max(time.perf_counter() - t0, 1e-3)
    nt = int(min(sysic.ntp, max(ncal, ncal * (0.7 * target_s) / tcal)))
    t0 = time.perf_counter()
    O.accel_pltp(sysic.tp_rh[:nt], rpl, Gm, omp=True)
    t_tp = time.perf_counter() - t0
    pairs = float(nt) * npl
    # pl-pl rows: a contiguous block of rows against all bodies
    t0 = time.perf_counter()
    O.lib(True).swo_kick_getacch_int_all_tri(npl, min(npl, 4096), np.ascontiguousarray(rpl), np.ascontiguousarray(Gm),
                                             None, np.zeros((npl, 3)))
    t_pp = time.perf_counter() - t0
    pairs_pp = float(min(npl, 4096)) * (npl - 1)
    rate = (pairs + pairs_pp) / (t_tp + t_pp)
    return {"value": rate, "unit": "pair-interactions/s", "cores": cores, "kind": "port", "cpu": cpu_model(),
            "sample": (f"CPU restatement (C, -O3, OpenMP, {cores} threads, reference loop decomposition) pair "
                       f"kernels on the benchmark's initial state: pl-tp accelerations of {nt} particles x {npl} "
                       f"bodies ({t_tp:.2f} s) plus 4096 full pl-pl rows ({t_pp:.2f} s); kernel time only. A whole "
                       f"step of this workload on the CPU is bound by its serial recursion and did not finish in "
                       f"25 minutes (one core); whole-step rates of the small workloads are under small_workloads")}


def small_workloads_leg(log, cpu=True, c2_cpu_steps_1thread=100):
    """Whole runs of the two small workloads, on the GPU and on the host:
    WHM 8 planets + 1000 particles (dt 0.01, 10^4 steps) and the SyMBA
    planetesimal disk of 2000 bodies (dt 0.005, 10^3 steps, CHK_CLOSE on),
    both in the exact pair mode (bit-identical to the CPU restatement).  The
    CPU side is the restatement's OpenMP build at the host's threads and at
    one thread (the disk's one-thread run over its first 100 steps)."""
    from swiftest_amd import ics
    import swiftest_amd as S
    out = {}
    cases = [("C1_whm_8pl_1000tp", S.WHM, ics.whm_solar(), 0.01, 10000, True),
             ("C2_symba_disk_2000pl", S.SYMBA, ics.symba_disk(), 0.005, 1000, True)]
    for name, integ, sysic, dt, nsteps, lclose in cases:
        g = S.GpuSystem(integ, sysic, dt, istep_out=100, lclose=lclose)
        g.run(1)
        t0 = time.perf_counter()
        g.run(nsteps - 1)
        t_gpu = time.perf_counter() - t0
        res = {"npl": sysic.npl, "ntp": sysic.ntp, "steps": nsteps, "dt": dt,
               "gpu": {"steps_per_s": (nsteps - 1) / t_gpu, "seconds": t_gpu,
                       "how": "one system, exact pair mode, steps 2..N timed"}}
        del g
        log(f"[bench] {name}: GPU {res['gpu']['steps_per_s']:.1f} steps/s")
        if integ == S.SYMBA and lclose and sysic.ntp == 0:
            # the workload's merge-only collisions resolved inside the step
            # (swg_system_set_merge): a step that collides runs again from
            # its start with a stop at each substep's collision check
            g = S.GpuSystem(integ, sysic, dt, istep_out=100, lclose=lclose, merge=(ics.GU, None))
            g.run(1)
            t0 = time.perf_counter()
            g.run(nsteps - 1)
            t_m = time.perf_counter() - t0
            res["gpu_merge_mode"] = {"steps_per_s": (nsteps - 1) / t_m, "seconds": t_m, "merges": g.nmerge}
            del g
            log(f"[bench] {name}: GPU merge mode {res['gpu_merge_mode']['steps_per_s']:.1f} steps/s")
        if cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as O
            oi = {S.WHM: O.WHM, S.HELIO: O.HELIO, S.SYMBA: O.SYMBA}[integ]
            for nth in (cpu_threads(), 1):
                steps = nsteps if (nth > 1 or integ == S.WHM) else c2_cpu_steps_1thread
                o = O.OracleSystem(oi, sysic, dt, istep_out=100, lclose=lclose, omp=True, nthreads=nth)
                t0 = time.perf_counter()
                rc = o.run(steps)
                t_cpu = time.perf_counter() - t0
                del o
                res[f"cpu_{nth}_threads"] = {"steps_per_s": steps / t_cpu, "seconds": t_cpu, "steps": steps, "rc": rc}
                log(f"[bench] {name}: CPU {nth} threads {steps / t_cpu:.1f} steps/s ({steps} steps)")
            # creating a system sets the OpenMP thread count of the process;
            # put it back for the legs that follow
            O.OracleSystem(oi, ics.whm_solar(npl=2, ntp=0), dt, omp=True, nthreads=cpu_threads())
            res["cpu"] = {"model": cpu_model(), "kind": "port",
                          "how": "CPU restatement, -O3 -fopenmp, reference loop decomposition, whole steps"}
            res["gpu_over_cpu"] = res["gpu"]["steps_per_s"] / res[f"cpu_{cpu_threads()}_threads"]["steps_per_s"]
            res["gpu_over_cpu_1thread"] = res["gpu"]["steps_per_s"] / res["cpu_1_threads"]["steps_per_s"]
        out[name] = res
    return out


def detection_leg(sysic, log, sparse_npl=131072):
    """Level-0 pl-pl encounter check with the triangular all-pairs sweep and
    with the sort-and-sweep broad phase (same list), on the benchmark's
    initial state and on a large planetesimal disk (C2 mass law, few pairs)."""
    from swiftest_amd import ics
    import swiftest_amd as S
    out = {}
    cases = [("c3", sysic), ("planetesimal_disk", ics.symba_disk(seed=2, npl=sparse_npl))]
    for label, case in cases:
        res = {"npl": case.npl}
        for mode, name in ((0, "triangular"), (1, "sortsweep")):
            g = S.GpuSystem(S.SYMBA, case, 0.005, istep_out=100, lclose=False, encounter_check=mode)
            ms, npairs, ncand = g.time_encounters(3)
            del g
            res[name] = {"ms": ms, "pairs": npairs}
            if ncand >= 0:
                res[name]["candidates"] = ncand
            log(f"[bench] {label} pl-pl encounter check {name}: {ms:.3f} ms, {npairs} pairs"
                + (f", {ncand} candidates" if ncand >= 0 else ""))
        n = case.npl
        res["pair_checks_per_s_triangular"] = n * (n - 1) / 2 / (res["triangular"]["ms"] * 1e-3)
        out[label] = res
    return out


def output_leg(g, sysic, log):
    """The output-cadence work on the benchmark system after its timed steps:
    energy and angular momentum (one O(N^2) pair-potential sum) and the
    discard checks with the thresholds of the reference's 108-body example
    (CHK_RMIN = CHK_QMIN = solar radius, CHK_RMAX = CHK_EJECT = 1000 AU,
    HELIO pericentres).  Each is timed over three calls after a first one
    that allocates, host round trip included."""
    import time
    reps = 3
    g.energy()
    t0 = time.perf_counter()
    for _ in range(reps):
        e = g.energy()
    ms_e = 1e3 * (time.perf_counter() - t0) / reps
    kw = dict(rmin=sysic.Rcb, rmax=1000.0, rmaxu=1000.0, qmin=sysic.Rcb, qmin_alo=sysic.Rcb, qmin_ahi=1000.0,
              qmin_coord=0)
    first = g.discard(**kw)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.discard(**kw)
    ms_d = 1e3 * (time.perf_counter() - t0) / reps
    n = sysic.npl
    res = {"energy_ms": ms_e, "energy_pairs_per_s": n * (n - 1) / 2 / (ms_e * 1e-3), "te": float(e[0, 5]),
           "discard_ms": ms_d, "discarded_first_call": list(first), "discard_bodies": n + sysic.ntp}
    log(f"[bench] energy {ms_e:.3f} ms ({res['energy_pairs_per_s']:.3e} pairs/s), discard checks {ms_d:.3f} ms "
        f"(first call discarded {first})")
    return res


def ensemble_leg(nsys, steps, warmup, sample, log):
    """Batched ensemble (config C4): nsys independent SyMBA systems of uneven
    size advanced together, one launch per step phase for all of them, with
    the pair kernels' tiles picked by the autotuner.  Compared with the same
    systems run one at a time through the same library (the plain GPU path),
    measured on a seeded sample of `sample` systems and scaled."""
    import numpy as np
    from swiftest_amd import ics
    import swiftest_amd as S
    dt = 0.01
    systems = ics.ensemble(seed=4, nsys=nsys)
    g = S.GpuSystem(S.SYMBA, systems, dt, istep_out=0, lclose=False)
    blocks = g.autotune(3)
    g.use_graphs(True)
    g.run(warmup)
    t0 = time.perf_counter()
    g.run(steps)
    t_ens = time.perf_counter() - t0
    st = g.stats()
    del g
    rate = nsys * steps / t_ens
    log(f"[bench] ensemble {nsys} systems x {steps} steps in {t_ens:.3f} s -> {rate:.1f} system-steps/s "
        f"(tiles {blocks}, last step {st})")
    rng = np.random.Generator(np.random.PCG64(44))
    idx = rng.choice(nsys, size=min(sample, nsys), replace=False)
    t_seq = 0.0
    for i in idx:
        one = S.GpuSystem(S.SYMBA, [systems[int(i)]], dt, istep_out=0, lclose=False)
        one.use_graphs(True)
        one.run(warmup)
        t0 = time.perf_counter()
        one.run(steps)
        t_seq += time.perf_counter() - t0
        del one
    seq_rate = len(idx) * steps / t_seq
    log(f"[bench] sequential sample {len(idx)} systems -> {seq_rate:.1f} system-steps/s")
    return {"metric": "ensemble system-steps/s", "value": rate, "unit": "system-steps/s",
            "systems": nsys, "steps": steps, "warmup": warmup, "seconds": t_ens,
            "npl_total": int(sum(s.npl for s in systems)), "ntp_total": int(sum(s.ntp for s in systems)),
            "tiles": {"kick_pl_block": blocks[0], "kick_tp_block": blocks[1]}, "hip_graphs": True,
            "sequential": {"value": seq_rate, "sample": int(len(idx)),
                           "how": "the same library, one system per handle, systems run one after another"},
            "speedup_vs_sequential": rate / seq_rate,
            "config": "SyMBA, dt 0.01 yr, N log-uniform 4-512, m log-uniform 1e-8..1e-4 Msun, a 0.3-30 AU, "
                      "ntp U{0..2000}, seed 4, CHK_CLOSE off"}


def serving_leg(nreq, rate, nmax, tpmax, smin, smax, chunk, seq_sample, log):
    """Served stream (config C5): C3-law systems (m log-uniform 1e-7..1e-3
    Msun over a U(0.5,5) AU, test particles a U(0.3,10) AU) with N
    log-uniform 8..nmax, U{0..tpmax} test particles and log-uniform
    smin..smax steps, arriving as a Poisson stream at `rate` per second,
    served with continuous batching over slot pools (swiftest_amd.serve).
    Compared with a seeded sample of the same requests run one after another
    through the plain path, one system per handle."""
    import numpy as np
    from swiftest_amd import ics, serve
    import swiftest_amd as S
    rng = np.random.Generator(np.random.PCG64(5))
    n = np.floor(np.exp(rng.uniform(np.log(8), np.log(nmax + 1), nreq))).astype(int)
    m = rng.integers(0, tpmax + 1, nreq)
    steps = np.floor(np.exp(rng.uniform(np.log(smin), np.log(smax + 1), nreq))).astype(int)
    systems = [ics.symba_big(seed=1000 + k, npl=int(n[k]), ntp=int(m[k])) for k in range(nreq)]
    reqs = serve.poisson_requests(systems, steps, rate, seed=6)
    srv = serve.Server(dt=0.005, chunk=chunk)
    wall = srv.run(reqs)
    srv.close()
    out = serve.summary(reqs, wall)
    occ = [float(np.mean(o)) if o else 0.0 for o in srv.occupancy]
    out.update({"metric": "served completion latency p50/p99 and system-steps/s", "arrival_rate_per_s": rate,
                "slot_classes": [{"npl_cap": c.npl_cap, "ntp_cap": c.ntp_cap, "slots": c.slots, "pools": c.pools,
                                  "loads": srv.loads[k], "mean_occupied": occ[k], "bytes": srv.pool_bytes[k]}
                                 for k, c in enumerate(srv.classes)],
                "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                "chunk_steps": chunk,
                "config": (f"SyMBA dt 0.005, C3 mass/orbit laws, N log-uniform 8-{nmax}, ntp U{{0..{tpmax}}}, "
                           f"steps log-uniform {smin}-{smax} (the law's 1e3-1e5 scaled down to keep the run "
                           f"short), {nreq} Poisson requests at {rate}/s, seed 5, CHK_CLOSE off, exact pair mode")})
    log(f"[bench] served {nreq} requests in {wall:.2f} s: p50 {out['p50_s']:.3f} s, p99 {out['p99_s']:.3f} s, "
        f"{out['system_steps_per_s']:.0f} system-steps/s, pools {out['slot_classes']}")
    # the same requests, one after another, each alone
    # every request (seq_sample <= 0 or >= nreq), or a seeded sample of them
    if seq_sample <= 0 or seq_sample >= nreq:
        idx = np.arange(nreq)
    else:
        idx = np.random.Generator(np.random.PCG64(55)).choice(nreq, size=seq_sample, replace=False)
    t_seq, s_seq = 0.0, 0
    for i in idx:
        r = reqs[int(i)]
        t0 = time.perf_counter()
        one = S.GpuSystem(S.SYMBA, [r.system], 0.005, istep_out=0, lclose=False)
        one.run(r.nsteps)
        t_seq += time.perf_counter() - t0
        s_seq += r.nsteps
        del one
        log(f"[bench] alone: request {int(i)} (N {r.system.npl}, {r.system.ntp} particles, {r.nsteps} steps) "
            f"{time.perf_counter() - t0:.2f} s")
    out["sequential"] = {"system_steps_per_s": s_seq / t_seq, "sample": int(len(idx)), "seconds": t_seq,
                         "how": ("the same requests" if len(idx) == nreq else "a sample of the same requests")
                         + ", each created and run alone through the plain path, one after another"}
    out["speedup_vs_sequential"] = out["system_steps_per_s"] / out["sequential"]["system_steps_per_s"]
    log(f"[bench] sequential sample {len(idx)} requests -> {s_seq / t_seq:.1f} system-steps/s")
    return out


def no_short_circuit_leg(sysic, mode, log, steps=1):
    """One step of the benchmark workload with the pl-tp encounter check run
    in every step (short_circuit off: the other legal evaluation of the
    reference's `pl check .or. tp check`, symba_step.f90:38), so that the
    pl-tp detection over npl x ntp pairs and the pl-tp recursion are timed,
    with their encounter counts."""
    import swiftest_amd as S
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, istep_out=100, lclose=False, short_circuit=False, pair_mode=mode)
    g.autotune(1)
    res = {"steps": steps, "step_seconds": [], "stats": []}
    for _ in range(steps):
        t0 = time.perf_counter()
        g.run(1)
        res["step_seconds"].append(time.perf_counter() - t0)
        res["stats"].append(g.stats())
        log(f"[bench] short_circuit off: step {res['step_seconds'][-1]:.3f} s {res['stats'][-1]}")
    del g
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--npl", type=int, default=32768)
    ap.add_argument("--ntp", type=int, default=1048576)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--istep-out", type=int, default=100)
    ap.add_argument("--lclose", type=int, default=0,
                    help="CHK_CLOSE: 1 checks collisions inside the recursion (their merge is host work, "
                         "so a detected collision stops the run)")
    ap.add_argument("--pair-mode", choices=("fast", "exact"), default="fast",
                    help="pair factor of the direct sums in the timed steps")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the whole-run legs of the two small workloads")
    ap.add_argument("--no-detection", action="store_true", help="skip the encounter-detection leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ens-systems", type=int, default=4096)
    ap.add_argument("--ens-steps", type=int, default=100)
    ap.add_argument("--ens-sample", type=int, default=32)
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-nmax", type=int, default=4096)
    ap.add_argument("--serve-tpmax", type=int, default=100000)
    ap.add_argument("--serve-steps", type=int, nargs=2, default=[20, 200])
    ap.add_argument("--serve-chunk", type=int, default=8)
    ap.add_argument("--serve-seq-sample", type=int, default=0,
                    help="requests run one after another for the comparison (0: all of them)")
    ap.add_argument("--no-sc-leg", action="store_true",
                    help="skip the step with the pl-tp encounter check not short-circuited")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the served-stream leg, printed as one JSON object (the full run starts it this way "
                         "in a process of its own)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 2**27 doubles (1 GiB) device to device three times: the known-size copy "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        x = torch.arange(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        assert torch.equal(x, y)
        print(json.dumps({"copy_calibration_bytes_read": x.numel() * 8, "bytes_written": y.numel() * 8,
                          "copies": 3}), flush=True)
        return 0

    # hardware queues for the served stream's pools, one stream each (read by
    # the HIP runtime when this process first initializes it, below)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path.", file=sys.stderr)
        return 2

    serve_args = (args.serve_requests, args.serve_rate, args.serve_nmax, args.serve_tpmax, args.serve_steps[0],
                  args.serve_steps[1], args.serve_chunk, args.serve_seq_sample)
    if args.serve_only:
        log = lambda *a: print(*a, file=sys.stderr, flush=True)
        print(json.dumps(serving_leg(*serve_args, log)), flush=True)
        return 0

    from swiftest_amd import ics
    import swiftest_amd as S

    sysic = ics.symba_big(seed=args.seed, npl=args.npl, ntp=args.ntp)
    dt = 0.005
    mode = S.PAIR_FAST if args.pair_mode == "fast" else S.PAIR_EXACT
    g = S.GpuSystem(S.SYMBA, sysic, dt, istep_out=args.istep_out, lclose=bool(args.lclose), pair_mode=mode)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    log(f"[bench] created system npl={sysic.npl} ntp={sysic.ntp}, pair mode {args.pair_mode}")
    tiles = g.autotune(2)
    log(f"[bench] autotuned tiles (kick_pl, kick_tp) = {tiles}")
    for k in range(args.warmup):
        t0 = time.perf_counter()
        g.run(1)
        log(f"[bench] warmup step {k} {time.perf_counter() - t0:.3f} s {g.stats()}")

    g.kernel_timing(True)
    per_step = []
    stats = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        g.run(1)  # returns after the step's work has completed on the device
        per_step.append(time.perf_counter() - t0)
        stats.append(g.stats())
        log(f"[bench] step {k} {per_step[-1]:.3f} s {stats[-1]}")
    elapsed = time.perf_counter() - t_start
    ms_tp, n_tp, ms_pl, n_pl = g.kernel_times()
    g.kernel_timing(False)

    npl, ntp = sysic.npl, sysic.ntp
    pairs_per_eval = npl * (npl - 1) / 2 + float(npl) * ntp
    tp_flops = FLOPS_PER_PLTP_PAIR * float(npl) * ntp
    pl_flops = FLOPS_PER_PLPL_PAIR * float(npl) * (npl - 1)
    value = 2 * pairs_per_eval * args.steps / elapsed
    achieved = tp_flops / (ms_tp * 1e-3) / 1e12 if ms_tp > 0 else 0.0
    pl_tflops = pl_flops / (ms_pl * 1e-3) / 1e12 if ms_pl > 0 else 0.0
    # both pair modes' kernels on the final state, same tiles policy (no
    # encounter-pair walks), for the side-by-side
    kmodes = {}
    for name, m in (("fast", S.PAIR_FAST), ("exact", S.PAIR_EXACT)):
        g.set_pair_mode(m)
        t = g.autotune(2)
        tms, pms = g.time_kernels(3)
        kmodes[name] = {"kick_tp_ms": tms, "kick_tp_frac": tp_flops / (tms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                        "kick_pl_ms": pms, "kick_pl_frac": pl_flops / (pms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                        "tiles": t}
        log(f"[bench] {name} pair kernels: kick_tp {tms:.3f} ms ({kmodes[name]['kick_tp_frac']:.3f} of peak), "
            f"kick_pl {pms:.3f} ms ({kmodes[name]['kick_pl_frac']:.3f} of peak), tiles {t}")
    g.set_pair_mode(mode)
    kname = "k_kick_tp_fast" if mode == S.PAIR_FAST else "k_kick_tp"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"SyMBA {npl} massive bodies + {ntp} test particles, dt={dt} yr, "
                                f"seed {args.seed}, snapshots every {args.istep_out} steps"),
                   "npl": npl, "ntp": ntp, "dt": dt, "integrator": "SyMBA", "short_circuit": True,
                   "chk_close": bool(args.lclose), "pair_mode": args.pair_mode},
        "steps_per_s": args.steps / elapsed,
        "step_seconds": per_step,
        "encounters": {"nplpl_enc": [s["nplpl_enc"] for s in stats], "npltp_enc": [s["npltp_enc"] for s in stats],
                       "max_depth": [s["max_depth"] for s in stats], "nsubsteps": [s["nsubsteps"] for s in stats]},
        "kernels": {"kick_tp_ms": ms_tp, "kick_tp_launches": n_tp, "kick_pl_ms": ms_pl, "kick_pl_launches": n_pl,
                    "kick_pl_tflops": pl_tflops, "kick_pl_frac": pl_tflops / PEAK_FP64_TFLOPS,
                    "kernel_pair_rate": pairs_per_eval / ((ms_tp + ms_pl) * 1e-3) if ms_tp + ms_pl > 0 else 0.0,
                    "by_pair_mode": kmodes},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(kname),
                     "kernel": f"{kname} (pl-tp accelerations, {args.pair_mode} pair factor)",
                     "note": ("FP64 vector-ALU-bound kernel (no MFMA: the pair sum is not a contraction); peak is "
                              "MI355X FP64 vector; 17 flops per pl-tp pair x npl x ntp per launch, duration from "
                              "HIP events around each launch of the timed steps; the fast factor takes 16 VALU "
                              "operations and one quarter-rate v_rsq_f64 per pair; traffic in bytes per launch "
                              "from rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by a copy calibration; the SyMBA step "
                              "itself is dominated by the recursion, reported under encounters")},
        "cpu_baseline": None,
    }
    out["output_cadence"] = output_leg(g, sysic, log)
    del g
    if not args.no_sc_leg:
        out["no_short_circuit"] = no_short_circuit_leg(sysic, mode, log)
    if not args.no_detection:
        out["encounter_detection"] = detection_leg(sysic, log)
    if not args.no_small:
        out["small_workloads"] = small_workloads_leg(log, cpu=not args.no_cpu)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ens_systems, args.ens_steps, 1, args.ens_sample, log)
    if not args.no_serve:
        # in a process of its own: the served stream's pools and threads get a
        # runtime that the legs above have not filled
        import subprocess
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serve-only",
               "--serve-requests", str(args.serve_requests), "--serve-rate", str(args.serve_rate),
               "--serve-nmax", str(args.serve_nmax), "--serve-tpmax", str(args.serve_tpmax),
               "--serve-steps", str(args.serve_steps[0]), str(args.serve_steps[1]),
               "--serve-chunk", str(args.serve_chunk), "--serve-seq-sample", str(args.serve_seq_sample)]
        env = dict(os.environ)
        if int(env.get("GPU_MAX_HW_QUEUES", "4")) < 32:
            env["GPU_MAX_HW_QUEUES"] = "32"  # one stream per pool; 15 pools by default
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"served-stream leg failed with status {r.returncode}")
        out["serving"] = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
        out["serving"]["process"] = "own"
    if not args.no_cpu:
        log("[bench] timing the CPU baseline sample")
        out["cpu_baseline"] = cpu_baseline(sysic, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
