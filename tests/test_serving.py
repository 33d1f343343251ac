"""Serving loop: admission and per-request completion (CPU, with a stand-in
for the device ensemble), and on the GPU each served result equals the same
request run alone."""
import threading

import numpy as np
import pytest

import swiftest_amd as S
from swiftest_amd import core, ics, serve


class FakeEnsemble:
    """Counts steps per system the way the library's freeze mask would."""
    live = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, integrator, systems, dt, lclose=False, istep_out=0, ctx=None):
        self.n = len(systems)
        self.steps = np.zeros(self.n, np.int64)
        self.frozen = np.zeros(self.n, np.int32)
        self.npl = [s.npl for s in systems]
        self.ntp = [s.ntp for s in systems]
        with FakeEnsemble.lock:
            FakeEnsemble.live += 1
            FakeEnsemble.peak = max(FakeEnsemble.peak, FakeEnsemble.live)

    def run(self, k):
        import time
        time.sleep(0.002 * k)
        self.steps[self.frozen == 0] += k

    def set_frozen(self, f):
        self.frozen = np.asarray(f, np.int32).copy()

    def state(self):
        npl, ntp = sum(self.npl), sum(self.ntp)
        # encode each system's step count in its bodies' positions
        rep_pl = np.repeat(self.steps, self.npl).astype(float)
        rep_tp = np.repeat(self.steps, self.ntp).astype(float)
        z = lambda n: np.zeros((n, 3))
        return {"pl_rh": np.stack([rep_pl] * 3, 1), "pl_vh": z(npl), "pl_id": np.arange(1, npl + 1, dtype=np.int32),
                "pl_status": np.zeros(npl, np.int32), "tp_rh": np.stack([rep_tp] * 3, 1), "tp_vh": z(ntp),
                "tp_status": np.zeros(ntp, np.int32)}

    def __del__(self):
        with FakeEnsemble.lock:
            FakeEnsemble.live -= 1


def test_admission_budget_and_own_step_counts(monkeypatch):
    monkeypatch.setattr(core, "new_context", lambda device=0: None)
    monkeypatch.setattr(core, "GpuSystem", FakeEnsemble)
    systems = ics.ensemble(seed=9, nsys=40, nmin=4, nmax=30, tpmax=20)
    rng = np.random.Generator(np.random.PCG64(2))
    nsteps = rng.integers(1, 40, len(systems))
    reqs = serve.poisson_requests(systems, nsteps, rate=1e9, seed=3)  # all arrive at once
    budget = 10 * max(serve.request_bytes(s) for s in systems)  # room for about two batches
    srv = serve.Server(lanes=3, batch_max=5, hbm_budget=budget)
    used_peak = [0]
    orig = srv._release

    def release(mem):
        used_peak[0] = max(used_peak[0], srv._used)
        orig(mem)
    srv._release = release
    wall = srv.run(reqs)
    assert wall > 0
    assert used_peak[0] <= budget
    for r in reqs:
        assert r.done >= r.admitted >= r.arrival - 1e-9
        # the stand-in writes each system's own step count into its positions
        if r.system.npl:
            assert np.all(r.state["pl_rh"] == r.nsteps)
    assert FakeEnsemble.peak >= 2          # batches overlapped
    assert max(n for _, n in srv.batches) > 1
    s = serve.summary(reqs, wall)
    assert s["p99_s"] >= s["p50_s"] > 0


@pytest.mark.gpu
def test_served_results_equal_each_request_alone():
    systems = ics.ensemble(seed=7, nsys=10, nmin=4, nmax=40, tpmax=60)
    nsteps = [5, 17, 9, 30, 12, 7, 25, 3, 14, 20]
    dt = 0.005
    reqs = serve.poisson_requests(systems, nsteps, rate=2000.0, seed=1)
    srv = serve.Server(dt=dt, lanes=3, batch_max=4)
    srv.run(reqs)
    srv.close()
    assert max(n for _, n in srv.batches) > 1
    for r in reqs:
        print(f"request {r.rid}: npl={r.system.npl} ntp={r.system.ntp} steps={r.nsteps}", flush=True)
        alone = S.GpuSystem(S.SYMBA, [r.system], dt, lclose=False)
        alone.run(r.nsteps)
        a = alone.state()
        for k in ("pl_rh", "pl_vh", "pl_status", "tp_rh", "tp_vh", "tp_status"):
            assert np.array_equal(a[k], r.state[k]), (r.rid, k)
        assert np.array_equal(a["pl_id"], r.state["pl_id"]), r.rid
