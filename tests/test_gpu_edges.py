"""Edge cases of the boundary and of ensembles, on the GPU: empty inputs to
every leaf operator (nothing written, no error), a single body, and
ensembles that mix systems of one massive body, of two, with and without
particles, each equal to the oracle's run of that system alone."""
import numpy as np
import pytest

import pyoracle as O
import swiftest_amd as S
from helpers import compact_disk
from swiftest_amd import ics

pytestmark = pytest.mark.gpu


def test_empty_inputs_to_the_leaf_operators():
    z3 = np.zeros((0, 3))
    acc = S.swiftest_kick_getacch_int_all_tri(0, 0, z3, np.zeros(0), np.zeros(0), z3)
    assert acc.shape == (0, 3)
    acc = S.swiftest_kick_getacch_int_all_flat(0, np.zeros((0, 2), np.int32), z3, np.zeros(0), None, z3)
    assert acc.shape == (0, 3)
    r1 = np.array([[1.0, 0.0, 0.0]])
    acc = S.swiftest_kick_getacch_int_all_tp(0, 1, z3, r1, np.array([1e-5]), np.zeros(0, np.int32), z3)
    assert acc.shape == (0, 3)
    acc = S.swiftest_kick_getacch_int_all_tp(3, 0, np.ones((3, 3)), z3, np.zeros(0), np.ones(3, np.int32),
                                             np.full((3, 3), 0.25))
    assert np.array_equal(acc, np.full((3, 3), 0.25))  # no massive body: nothing added
    x, v, fl = S.swiftest_drift_all(np.zeros(0), z3, z3, 0, 0.01, np.zeros(0, np.int32))
    assert x.shape == (0, 3) and fl.shape == (0,)
    i1, i2, lv = S.encounter_check_all_plpl(0, z3, z3, np.zeros(0), 0.01)
    assert len(i1) == 0
    i1, i2, lv = S.encounter_check_all_plpl(1, r1, r1, np.ones(1), 0.01)
    assert len(i1) == 0
    i1, i2, lv = S.encounter_check_all_pltp(1, 0, r1, r1, z3, z3, np.ones(1), 0.01)
    assert len(i1) == 0


def test_single_body_accelerations_are_zero():
    acc = S.swiftest_kick_getacch_int_all_tri(1, 1, np.array([[1.0, 2.0, 3.0]]), np.array([1e-4]), np.array([1e-3]),
                                              np.zeros((1, 3)))
    assert np.array_equal(acc, np.zeros((1, 3)))


@pytest.mark.parametrize("integ", [S.SYMBA, S.HELIO])
def test_ensemble_of_degenerate_systems_equals_the_oracle(integ):
    systems = [ics.whm_solar(seed=90, npl=1, ntp=50), ics.whm_solar(seed=91, npl=2, ntp=0),
               compact_disk(seed=92, npl=30, ntp=1), ics.whm_solar(seed=93, npl=1, ntp=0),
               compact_disk(seed=94, npl=70, ntp=65)]
    g = S.GpuSystem(integ, systems, 0.005, istep_out=6)
    oracles = [O.OracleSystem(integ, s, 0.005, istep_out=6) for s in systems]
    off_pl = np.cumsum([0] + [s.npl for s in systems])
    off_tp = np.cumsum([0] + [s.ntp for s in systems])
    for blk in range(3):
        print(f"degenerate ensemble integ={integ} block {blk}", flush=True)
        g.run(10)
        sg = g.state()
        for k, o in enumerate(oracles):
            assert o.run(10) == 0
            so = o.state()
            for key in ("pl_rh", "pl_vh", "pl_vb"):
                assert np.array_equal(so[key], sg[key][off_pl[k]:off_pl[k + 1]]), (blk, k, key)
            for key in ("tp_rh", "tp_vh", "tp_vb"):
                assert np.array_equal(so[key], sg[key][off_tp[k]:off_tp[k + 1]]), (blk, k, key)
