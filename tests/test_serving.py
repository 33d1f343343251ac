"""Serving loop: continuous batching over slot pools.

On the CPU, with a stand-in for the device pools: every request runs exactly
its own step count, requests are admitted into free slots while other
requests of the same pool are running, and the pools stay within the memory
budget.  On the GPU: requests drawn with the served-stream law (N
log-uniform over 8..4096, test particles uniform over 0..100000, one of them
at N = 4096 and 100000 particles; step counts scaled down so that the oracle
finishes) are served in overlapping slots, and each result equals the
oracle's run of that request alone, bit for bit."""
import threading

import numpy as np
import pytest

import swiftest_amd as S
from swiftest_amd import core, ics, serve


class FakePool:
    """Counts steps per slot the way the library's freeze mask would."""
    live = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, npl_cap, ntp_cap):
        self.n = len(npl_cap)
        self.steps = np.zeros(self.n, np.int64)
        self.frozen = np.ones(self.n, np.int32)
        self.held = [None] * self.n
        self.cap = (list(npl_cap), list(ntp_cap))
        with FakePool.lock:
            FakePool.live += 1
            FakePool.peak = max(FakePool.peak, FakePool.live)

    def load(self, slot, system):
        assert system.npl <= self.cap[0][slot] and system.ntp <= self.cap[1][slot]
        assert self.frozen[slot] == 1  # only free slots take a new system
        self.held[slot] = system
        self.steps[slot] = 0
        self.frozen[slot] = 0

    def run(self, k):
        import time
        time.sleep(0.001 * k)
        self.steps[self.frozen == 0] += k

    def set_frozen(self, f):
        self.frozen = np.asarray(f, np.int32).copy()

    def get_slot(self, slot):
        s = self.held[slot]
        return {"pl_rh": np.full((s.npl, 3), float(self.steps[slot])), "tp_rh": np.zeros((s.ntp, 3))}

    def __del__(self):
        with FakePool.lock:
            FakePool.live -= 1


class FakeGpuSystem:
    @classmethod
    def slots(cls, integrator, npl_cap, ntp_cap, dt, lclose=True, ctx=None, pair_mode=0):
        return FakePool(npl_cap, ntp_cap)


class FakeLib:
    def swg_finalize(self, ctx):
        return 0


def test_continuous_admission_and_own_step_counts(monkeypatch):
    monkeypatch.setattr(core, "new_context", lambda device=0: None)
    monkeypatch.setattr(core, "GpuSystem", FakeGpuSystem)
    monkeypatch.setattr(core, "library", lambda: FakeLib())
    systems = ics.ensemble(seed=9, nsys=60, nmin=4, nmax=300, tpmax=20)
    rng = np.random.Generator(np.random.PCG64(2))
    nsteps = rng.integers(1, 60, len(systems))
    reqs = serve.poisson_requests(systems, nsteps, rate=2000.0, seed=3)
    classes = [serve.SlotClass(32, 64, 3), serve.SlotClass(512, 64, 2)]
    budget = sum(c.slots * serve.request_bytes(c.npl_cap, c.ntp_cap) for c in classes)
    srv = serve.Server(classes=classes, hbm_budget=budget, chunk=4)
    wall = srv.run(reqs)
    assert wall > 0
    assert FakePool.peak == 2               # both pools ran at the same time
    for r in reqs:
        assert r.done >= r.admitted >= r.arrival - 1e-9
        # the stand-in writes each slot's own step count into its positions
        if r.system.npl:
            assert np.all(r.state["pl_rh"] == r.nsteps), r.rid
    # slot-level admission: a request entered a pool while another request
    # of that pool was already running and not yet done
    overlap = any(a.lane == b.lane and a.admitted < b.admitted < a.done for a in reqs for b in reqs if a is not b)
    assert overlap
    assert max(max(o) for o in srv.occupancy if o) > 1
    s = serve.summary(reqs, wall)
    assert s["p99_s"] >= s["p50_s"] > 0
    with pytest.raises(ValueError):
        serve.Server(classes=classes, hbm_budget=budget - 1)


def _c5_requests(nreq, seed):
    """Systems drawn with the served-stream law; the first is the largest
    request the law allows.  Step counts log-uniform over 1..6 (the law's
    1e3..1e5 scaled down so that the oracle runs finish)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = np.floor(np.exp(rng.uniform(np.log(8), np.log(4097), nreq))).astype(int)
    m = rng.integers(0, 100001, nreq)
    steps = np.floor(np.exp(rng.uniform(0, np.log(7), nreq))).astype(int)
    n[0], m[0], steps[0] = 4096, 100000, 1
    return [ics.symba_big(seed=5000 + k, npl=int(n[k]), ntp=int(m[k])) for k in range(nreq)], steps


@pytest.mark.gpu
def test_served_results_equal_the_oracle_alone():
    import pyoracle as O
    systems, nsteps = _c5_requests(16, seed=8)
    dt = 0.005
    reqs = serve.poisson_requests(systems, nsteps, rate=200.0, seed=1)
    classes = [serve.SlotClass(128, 100000, 4), serve.SlotClass(1024, 100000, 3), serve.SlotClass(4096, 100000, 2)]
    srv = serve.Server(classes=classes, dt=dt, chunk=1)
    srv.run(reqs)
    srv.close()
    assert any(a.lane == b.lane and a.admitted < b.admitted < a.done for a in reqs for b in reqs if a is not b)
    for r in reqs:
        print(f"request {r.rid}: npl={r.system.npl} ntp={r.system.ntp} steps={r.nsteps} lane={r.lane} "
              f"slot={r.slot}", flush=True)
        o = O.OracleSystem(O.SYMBA, r.system, dt, lclose=False, omp=True, nthreads=16)
        assert o.run(r.nsteps) == 0
        so = o.state()
        for k in ("pl_rh", "pl_vh", "pl_vb", "tp_rh", "tp_vh", "tp_vb", "pl_status", "tp_status"):
            assert np.array_equal(so[k], r.state[k]), (r.rid, k)
        assert np.array_equal(so["pl_id"], r.state["pl_id"]), r.rid

