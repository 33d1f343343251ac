"""Serving loop for integration requests.

Requests (one planetary system and a step count each) arrive over time.  A
request is admitted once it has arrived, a lane is free and the device-memory
budget has room for it; the requests admitted together form one batch, an
ensemble advanced by one launch per step phase.  Each lane is a host thread
with its own library context (its own HIP stream), so the batches of
different lanes run on the device at the same time, and requests that arrive
while other batches are running are admitted as lanes and memory free up.

Inside a batch every system stops at its own step count: the batch advances
to the next completion, the finished systems are frozen (left out of the
following steps with their state kept exactly), and the rest continue.  A
system's trajectory does not depend on the other systems of its batch, so
each result equals that of the request run alone.
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
    lane: int = -1
    state: dict = field(default_factory=dict)


def request_bytes(system) -> int:
    """Device memory a request holds while resident: body state, the
    encounter bit matrices (pl-pl and pl-tp, words plus prefixes) and pair
    lists sized at a few pairs per body."""
    n, m = system.npl, system.ntp
    words_pp = n * ((n + 63) // 64)
    words_pt = n * ((m + 63) // 64)
    return int(200 * n + 100 * m + 12 * (words_pp + words_pt) + 64 * (n + m) + (1 << 16))


class Server:
    """Admission by memory budget, `lanes` concurrent batches of at most
    `batch_max` requests, SyMBA (or HELIO) with one dt for all requests."""

    def __init__(self, integrator=core.SYMBA, dt=0.005, lanes=4, batch_max=128, hbm_budget=64 << 30,
                 lclose=False, istep_out=0):
        self.integrator, self.dt = integrator, dt
        self.lanes, self.batch_max, self.budget = lanes, batch_max, hbm_budget
        self.lclose, self.istep_out = lclose, istep_out
        self.ctx = [core.new_context() for _ in range(lanes)]
        self._lock = threading.Condition()
        self._used = 0
        self.batches = []

    # ---- admission ----
    def _take_batch(self, pending, t0):
        """Block until a batch can be formed; None when nothing is left."""
        with self._lock:
            while True:
                if not pending:
                    return None
                now = time.perf_counter() - t0
                batch, mem = [], 0
                for r in list(pending):
                    if r.arrival > now or len(batch) >= self.batch_max:
                        break
                    need = request_bytes(r.system)
                    if self._used + mem + need > self.budget and (batch or self._used):
                        break
                    batch.append(r)
                    mem += need
                if batch:
                    for r in batch:
                        pending.remove(r)
                        r.admitted = now
                    self._used += mem
                    return batch, mem
                # wait for the next arrival or for memory to be released
                wait = max(0.0, pending[0].arrival - now) if pending[0].arrival > now else 0.05
                self._lock.wait(timeout=min(wait, 0.05) if wait > 0 else 0.05)

    def _release(self, mem):
        with self._lock:
            self._used -= mem
            self._lock.notify_all()

    # ---- one lane ----
    def _lane(self, lane, pending, t0, errors):
        try:
            while True:
                got = self._take_batch(pending, t0)
                if got is None:
                    return
                batch, mem = got
                self._run_batch(lane, batch, t0)
                self._release(mem)
        except Exception as e:  # reported by run()
            errors.append(e)
            with self._lock:
                self._lock.notify_all()

    def _run_batch(self, lane, batch, t0):
        g = core.GpuSystem(self.integrator, [r.system for r in batch], self.dt, lclose=self.lclose,
                           istep_out=self.istep_out, ctx=self.ctx[lane])
        if hasattr(g, "use_graphs"):
            g.use_graphs(True)
        remaining = np.array([r.nsteps for r in batch], np.int64)
        frozen = np.zeros(len(batch), np.int32)
        while (remaining > 0).any():
            k = int(remaining[remaining > 0].min())
            g.run(k)
            remaining[remaining > 0] -= k
            now = time.perf_counter() - t0
            newly = (remaining == 0) & (frozen == 0)
            for q in np.flatnonzero(newly):
                batch[q].done = now
                batch[q].lane = lane
            frozen[newly] = 1
            if (remaining > 0).any():
                g.set_frozen(frozen)
        st = g.state()
        po = to = 0
        for r in batch:
            n, m = r.system.npl, r.system.ntp
            r.state = {
                "pl_rh": st["pl_rh"][po:po + n].copy(), "pl_vh": st["pl_vh"][po:po + n].copy(),
                "pl_id": st["pl_id"][po:po + n].copy(), "pl_status": st["pl_status"][po:po + n].copy(),
                "tp_rh": st["tp_rh"][to:to + m].copy(), "tp_vh": st["tp_vh"][to:to + m].copy(),
                "tp_status": st["tp_status"][to:to + m].copy()}
            po += n
            to += m
        self.batches.append((lane, len(batch)))
        del g

    # ---- driver ----
    def run(self, requests):
        """Serve `requests` (arrival times relative to the start); returns the
        wall time.  Each request gets admitted / done times and its state."""
        pending = sorted(requests, key=lambda r: r.arrival)
        errors = []
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._lane, args=(k, pending, t0, errors)) for k in range(self.lanes)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return time.perf_counter() - t0

    def close(self):
        for c in self.ctx:
            core.library().swg_finalize(c)
        self.ctx = []


def poisson_requests(systems, nsteps, rate, seed):
    """Requests for `systems` with Poisson arrivals at `rate` per second."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, len(systems)))
    return [Request(k, s, int(n), float(a)) for k, (s, n, a) in enumerate(zip(systems, nsteps, t))]


def summary(requests, wall):
    lat = np.array([r.done - r.arrival for r in requests])
    steps = sum(r.nsteps for r in requests)
    return {"requests": len(requests), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "system_steps_per_s": steps / wall, "wall_s": wall}
