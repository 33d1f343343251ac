"""Serving loop for integration requests, with continuous batching.

Requests (one planetary system and a step count each) arrive over time.  The
device holds slot pools (swg_system_create_slots): ensembles of fixed-capacity
slots allocated once, within the device-memory budget, one pool per size
class.  Each pool is driven by its own host thread with its own library
context (its own HIP stream), so the pools' steps overlap on the device.

A request goes to the smallest class its system fits.  A class may have
several pools, each with its own thread and stream, that take requests from
the class's queue: the SyMBA recursion of a pool's systems runs in lockstep
(a substep at some level is taken when any of its systems needs it), so
large crowded systems are better spread over several pools, whose steps
overlap, than packed into one.  Between steps a pool loads the requests that
have arrived into its free slots (swg_system_load) while the systems in its
other slots keep running; it then advances all of its occupied slots
together, one launch per step phase, for at most `chunk` steps or until the
next request in it completes.  A completed request's state
is copied out (swg_system_get_slot) and its slot is frozen and free for the
next arrival.  Padding bodies make a slot's system evolve exactly as it would
alone, so each result equals that of the request run by itself.

The pools are allocated when the server starts, before the clock of the
arrivals: arrival, admission and completion times count from the moment
every pool is ready.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import core


@dataclass
class Request:
    rid: int
    system: object          # swiftest_amd.ics.System
    nsteps: int
    arrival: float = 0.0    # seconds after the server starts
    admitted: float = -1.0
    done: float = -1.0
    lane: tuple = ()        # (class, pool) that served it
    slot: int = -1
    state: dict = field(default_factory=dict)


@dataclass
class SlotClass:
    """`pools` pools of `slots` slots each, every slot with room for npl_cap
    massive bodies and ntp_cap particles."""
    npl_cap: int
    ntp_cap: int
    slots: int
    pools: int = 1


def request_bytes(npl: int, ntp: int) -> int:
    """Device memory one slot of this capacity holds: body state, the
    encounter bit matrices (pl-pl and pl-tp, words plus prefixes) and pair
    lists sized at a few pairs per body."""
    n, m = npl, ntp
    words_pp = n * ((n + 63) // 64)
    words_pt = n * ((m + 63) // 64)
    return int(200 * n + 100 * m + 12 * (words_pp + words_pt) + 64 * (n + m) + (1 << 16))


# size classes for requests of up to 4096 massive bodies and 100000 particles:
# small systems packed 32 to a pool, the largest one per pool (their
# recursions are long and run best side by side on their own streams).
# The pools' streams overlap on the device only as far as the runtime gives
# the process hardware queues: set GPU_MAX_HW_QUEUES to at least the number
# of pools (14 here) before the process's first HIP call (bench.py uses 32).
DEFAULT_CLASSES = (SlotClass(128, 100000, 32, 1), SlotClass(1024, 100000, 4, 3), SlotClass(4096, 100000, 1, 10))


class Server:
    """Continuous batching over slot pools, SyMBA (or HELIO) with one dt for
    all requests."""

    def __init__(self, classes=DEFAULT_CLASSES, integrator=core.SYMBA, dt=0.005, hbm_budget=64 << 30,
                 lclose=False, chunk=8, pair_mode=core.PAIR_EXACT):
        self.classes = list(classes)
        self.integrator, self.dt, self.lclose, self.chunk = integrator, dt, lclose, max(1, int(chunk))
        self.pair_mode = pair_mode
        self.pool_bytes = [c.pools * c.slots * request_bytes(c.npl_cap, c.ntp_cap) for c in self.classes]
        if sum(self.pool_bytes) > hbm_budget:
            raise ValueError(f"slot pools need {sum(self.pool_bytes)} bytes, over the budget of {hbm_budget}")
        self.budget = hbm_budget
        self._lock = threading.Condition()
        self.loads = [0] * len(self.classes)      # requests loaded per pool
        self.occupancy = [[] for _ in self.classes]  # occupied slots at each chunk
        self.pools = []                             # (class index, pool) while running, for tests

    def class_of(self, system) -> int:
        for k, c in enumerate(self.classes):
            if system.npl <= c.npl_cap and system.ntp <= c.ntp_cap:
                return k
        raise ValueError(f"no slot class holds a system of {system.npl} + {system.ntp} bodies")

    # ---- one pool ----
    def _lane(self, k, p, pending, ready, errors):
        try:
            c = self.classes[k]
            ctx = core.new_context()
            try:
                pool = core.GpuSystem.slots(self.integrator, [c.npl_cap] * c.slots, [c.ntp_cap] * c.slots, self.dt,
                                            lclose=self.lclose, ctx=ctx, pair_mode=self.pair_mode)
                with self._lock:
                    self.pools.append((k, pool))
                ready.wait()  # every pool allocated: the clock starts
                self._drive(k, p, c, pool, pending, self._t0)
                del pool
            finally:
                core.library().swg_finalize(ctx)
        except Exception as e:  # reported by run()
            errors.append(e)
            ready.abort()
            with self._lock:
                self._lock.notify_all()

    def _drive(self, k, p, c, pool, pending, t0):
        held = [None] * c.slots            # the request each slot holds
        remaining = np.zeros(c.slots, np.int64)
        while True:
            now = time.perf_counter() - t0
            free = [s for s in range(c.slots) if held[s] is None]
            with self._lock:
                take = []
                while pending and len(take) < len(free) and pending[0].arrival <= now:
                    take.append(pending.pop(0))
                empty = not pending
                nxt = pending[0].arrival if pending else None
            for s, r in zip(free, take):
                pool.load(s, r.system)
                r.admitted = time.perf_counter() - t0
                r.lane, r.slot = (k, p), s
                held[s] = r
                remaining[s] = r.nsteps
                with self._lock:
                    self.loads[k] += 1
            busy = [s for s in range(c.slots) if held[s] is not None]
            if not busy:
                if empty:
                    return
                time.sleep(min(max(nxt - (time.perf_counter() - t0), 0.0), 0.05))
                continue
            self.occupancy[k].append(len(busy))
            n = int(min(self.chunk, remaining[busy].min()))
            pool.run(n)
            remaining[busy] -= n
            done = [s for s in busy if remaining[s] == 0]
            if done:
                t = time.perf_counter() - t0
                for s in done:
                    r = held[s]
                    r.state = pool.get_slot(s)
                    r.done = t
                    held[s] = None
                pool.set_frozen(np.array([h is None for h in held], np.int32))

    # ---- driver ----
    def run(self, requests):
        """Serve `requests` (arrival times relative to the start); returns the
        wall time.  Each request gets admitted / done times and its state."""
        queues = [[] for _ in self.classes]
        for r in sorted(requests, key=lambda r: r.arrival):
            queues[self.class_of(r.system)].append(r)
        errors = []
        lanes = [(k, p) for k in range(len(self.classes)) if queues[k] for p in range(self.classes[k].pools)]
        ready = threading.Barrier(len(lanes) + 1, action=lambda: setattr(self, "_t0", time.perf_counter()))
        threads = [threading.Thread(target=self._lane, args=(k, p, queues[k], ready, errors)) for k, p in lanes]
        for t in threads:
            t.start()
        try:
            ready.wait()
        except threading.BrokenBarrierError:
            pass  # a pool failed to start; its error is raised below
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return time.perf_counter() - self._t0

    def close(self):
        self.pools = []


def poisson_requests(systems, nsteps, rate, seed):
    """Requests for `systems` with Poisson arrivals at `rate` per second."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, len(systems)))
    return [Request(k, s, int(n), float(a)) for k, (s, n, a) in enumerate(zip(systems, nsteps, t))]


def summary(requests, wall):
    lat = np.array([r.done - r.arrival for r in requests])
    wait = np.array([r.admitted - r.arrival for r in requests])
    steps = sum(r.nsteps for r in requests)
    return {"requests": len(requests), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "admission_wait_p50_s": float(np.percentile(wait, 50)),
            "admission_wait_p99_s": float(np.percentile(wait, 99)), "system_steps_per_s": steps / wall,
            "wall_s": wall}
