"""Whole WHM steps in one launch (k_whm_run, systems of at most 64 massive
bodies) against the per-step kernels (SWG_WHM_STEPWISE=1) and the oracle,
bit for bit: an ensemble whose systems have 1, 8, 33 and 64 planets and
0, 1, 64, 65 and 300 particles, with snapshots splitting the runs into
launches of uneven length; single steps interleaved with whole-step
launches; the fast pair mode."""
import numpy as np
import pytest

import pyoracle as O
import swiftest_amd as S
from swiftest_amd import ics

pytestmark = pytest.mark.gpu

KEYS = ("pl_rh", "pl_vh", "tp_rh", "tp_vh", "pl_status", "tp_status")


def _systems():
    return [ics.whm_solar(seed=70 + k, npl=n, ntp=t)
            for k, (n, t) in enumerate(((8, 300), (1, 65), (33, 0), (64, 64), (8, 1)))]


def _equal(a, b, label):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (label, k)


def test_whole_step_launches_equal_per_step_kernels_and_the_oracle(monkeypatch):
    systems = _systems()
    fused = S.GpuSystem(S.WHM, systems, 0.01, istep_out=37)
    monkeypatch.setenv("SWG_WHM_STEPWISE", "1")
    stepwise = S.GpuSystem(S.WHM, systems, 0.01, istep_out=37)
    monkeypatch.delenv("SWG_WHM_STEPWISE")
    oracles = [O.OracleSystem(O.WHM, s, 0.01, istep_out=37) for s in systems]
    for blk, n in enumerate((1, 50, 200, 7)):
        print(f"whm whole-step block {blk}: {n} steps", flush=True)
        fused.run(n)
        stepwise.run(n)
        _equal(fused.state(), stepwise.state(), f"block {blk}")
        sf = fused.state()
        off_pl = np.cumsum([0] + [s.npl for s in systems])
        off_tp = np.cumsum([0] + [s.ntp for s in systems])
        for k, o in enumerate(oracles):
            assert o.run(n) == 0
            so = o.state()
            for key in ("pl_rh", "pl_vh"):
                assert np.array_equal(so[key], sf[key][off_pl[k]:off_pl[k + 1]]), (blk, k, key)
            for key in ("tp_rh", "tp_vh"):
                assert np.array_equal(so[key], sf[key][off_tp[k]:off_tp[k + 1]]), (blk, k, key)


def test_single_steps_between_whole_step_launches(monkeypatch):
    c1 = ics.whm_solar(seed=5, npl=8, ntp=200)
    fused = S.GpuSystem(S.WHM, c1, 0.01)
    monkeypatch.setenv("SWG_WHM_STEPWISE", "1")
    stepwise = S.GpuSystem(S.WHM, c1, 0.01)
    monkeypatch.delenv("SWG_WHM_STEPWISE")
    for n in (3, 1, 20, 1, 1, 9):
        if n == 1:
            fused.step()
        else:
            fused.run(n)
        stepwise.run(n)
        _equal(fused.state(), stepwise.state(), n)


def test_whole_step_launches_fast_pair_mode(monkeypatch):
    c1 = ics.whm_solar(seed=6, npl=8, ntp=500)
    fused = S.GpuSystem(S.WHM, c1, 0.01, pair_mode=S.PAIR_FAST)
    monkeypatch.setenv("SWG_WHM_STEPWISE", "1")
    stepwise = S.GpuSystem(S.WHM, c1, 0.01, pair_mode=S.PAIR_FAST)
    monkeypatch.delenv("SWG_WHM_STEPWISE")
    fused.run(300)
    stepwise.run(300)
    _equal(fused.state(), stepwise.state(), "fast")
