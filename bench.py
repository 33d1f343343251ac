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
and frame changes included).  The roofline object covers the dominant kernel
(pl-tp accelerations, 17 flops per pair) with its launch durations measured
by HIP events on its own stream during the timed steps.  The CPU baseline
times the CPU restatement's OpenMP pair kernels on a bounded sample of the
same state.
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


def cpu_baseline(sysic, target_s):
    """Pair-interaction rate of the CPU restatement's OpenMP kernels (the
    reference's loop decomposition: OpenMP over particles / rows, j in order)
    on a bounded sample of the benchmark state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O  # the checker doubles as the CPU baseline
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
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
    return {"value": rate, "unit": "pair-interactions/s", "cores": cores, "kind": "port",
            "sample": (f"CPU restatement (C, -O3, OpenMP, {cores} threads, reference loop decomposition) pair "
                       f"kernels on the benchmark's initial state: pl-tp accelerations of {nt} particles x {npl} "
                       f"bodies ({t_tp:.2f} s) plus 4096 full pl-pl rows ({t_pp:.2f} s); kernel time only, no "
                       f"encounter detection, recursion or drift")}


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
                    help="CHK_CLOSE: 1 checks collisions inside the recursion (their merge is host work "
                         "this build does not do, so a detected collision stops the run)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path.", file=sys.stderr)
        return 2

    from swiftest_amd import ics
    import swiftest_amd as S

    sysic = ics.symba_big(seed=args.seed, npl=args.npl, ntp=args.ntp)
    dt = 0.005
    g = S.GpuSystem(S.SYMBA, sysic, dt, istep_out=args.istep_out, lclose=bool(args.lclose))
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    log(f"[bench] created system npl={sysic.npl} ntp={sysic.ntp}")
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
    value = 2 * pairs_per_eval * args.steps / elapsed
    achieved = FLOPS_PER_PLTP_PAIR * float(npl) * ntp / (ms_tp * 1e-3) / 1e12 if ms_tp > 0 else 0.0
    pl_tflops = FLOPS_PER_PLPL_PAIR * float(npl) * (npl - 1) / (ms_pl * 1e-3) / 1e12 if ms_pl > 0 else 0.0
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
                   "chk_close": bool(args.lclose)},
        "steps_per_s": args.steps / elapsed,
        "encounters": {"nplpl_enc": [s["nplpl_enc"] for s in stats], "npltp_enc": [s["npltp_enc"] for s in stats],
                       "max_depth": [s["max_depth"] for s in stats], "nsubsteps": [s["nsubsteps"] for s in stats]},
        "kernels": {"kick_tp_ms": ms_tp, "kick_tp_launches": n_tp, "kick_pl_ms": ms_pl, "kick_pl_launches": n_pl,
                    "kick_pl_tflops": pl_tflops,
                    "kernel_pair_rate": pairs_per_eval / ((ms_tp + ms_pl) * 1e-3) if ms_tp + ms_pl > 0 else 0.0},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "k_kick_tp (pl-tp accelerations)",
                     "note": ("compute-bound FP64 VALU kernel, no MFMA; peak is MI355X FP64 vector; "
                              f"{FLOPS_PER_PLTP_PAIR} flops per pl-tp pair x npl x ntp per launch")},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("[bench] timing the CPU baseline sample")
        out["cpu_baseline"] = cpu_baseline(sysic, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
