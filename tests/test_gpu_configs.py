"""GPU against the oracle on the benchmark workloads themselves (the
generators in swiftest_amd.ics, at their full sizes):

C1  WHM, 8 planets + 1000 particles, dt 0.01, 10^4 steps: bit-exact;
C2  SyMBA planetesimal disk, 2000 bodies, dt 0.005, 10^3 steps, CHK_CLOSE
    on: bit-exact states, encounter lists, depth and substep counts at every
    100th step;
C3  32768 bodies + 1048576 particles: the level-0 pl-pl list, both
    direct-sum kernels (every massive body's row, a 65536-particle sample)
    and one Danby drift of all 1081344 bodies;
C4  the 4096-system ensemble, 10 steps, 64 sampled systems against their
    own single-system oracle runs;

and, for the fast pair mode (SWG_PAIR_FAST), the tolerances the benchmark's
correctness statement names: per-step accelerations within 1e-13 relative
(C1, C2, C3 states), positions within 1e-10 relative after 10^3 steps with the
same encounter lists (C1, C2).  Driver cadence: swiftest_driver.f90:130-180.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
import swiftest_amd as S
from swiftest_amd import ics

pytestmark = pytest.mark.gpu


def _rel(a, b):
    """largest |a - b| / |b| over bodies (3-vectors)"""
    d = np.linalg.norm(np.asarray(a) - np.asarray(b), axis=-1)
    n = np.linalg.norm(np.asarray(b), axis=-1)
    return float(np.max(d / np.where(n > 0, n, 1.0))) if d.size else 0.0


def _abs_terms_tp(rtp, rpl, gm):
    """sum over massive bodies of |Gm / r^2| for each particle (the scale of
    the sum's rounding)"""
    out = np.zeros(len(rtp))
    for k in range(0, len(rtp), 128):
        d = rtp[k:k + 128, None, :] - rpl[None, :, :]
        r2 = (d * d).sum(-1)
        out[k:k + 128] = (gm[None, :] / r2).sum(1)
    return out


# ------------------------------------------------------------------- C1 --

def test_c1_whm_1e4_steps_bit_exact():
    c1 = ics.whm_solar()
    o = O.OracleSystem(O.WHM, c1, 0.01, istep_out=1000)
    g = S.GpuSystem(S.WHM, c1, 0.01, istep_out=1000)
    for k in range(10):
        assert o.run(1000) == 0
        g.run(1000)
        so, sg = o.state(), g.state()
        for key in ("pl_rh", "pl_vh", "tp_rh", "tp_vh"):
            assert np.array_equal(so[key], sg[key]), (k, key)


def test_c1_fast_pairs_1e3_steps_within_1e_10():
    c1 = ics.whm_solar()
    o = O.OracleSystem(O.WHM, c1, 0.01, istep_out=100)
    g = S.GpuSystem(S.WHM, c1, 0.01, istep_out=100, pair_mode=S.PAIR_FAST)
    assert o.run(1000) == 0
    g.run(1000)
    so, sg = o.state(), g.state()
    assert _rel(sg["pl_rh"], so["pl_rh"]) <= 1e-10
    assert _rel(sg["tp_rh"], so["tp_rh"]) <= 1e-10


def test_c1_fast_accelerations_within_1e_13():
    c1 = ics.whm_solar()
    g = S.GpuSystem(S.HELIO, c1, 0.01, lclose=False, pair_mode=S.PAIR_FAST)
    e = S.GpuSystem(S.HELIO, c1, 0.01, lclose=False)
    st = e.state()
    gm = c1.Gmass[st["pl_id"] - 1]
    pf, tf = g.accel()
    pe, te = e.accel()
    assert np.array_equal(te, O.accel_pltp(st["tp_rh"], st["pl_rh"], gm))
    assert _rel(pf, pe) <= 1e-13
    assert _rel(tf, te) <= 1e-13


# ------------------------------------------------------------------- C2 --

@pytest.fixture(scope="module")
def c2_oracle():
    """The oracle's C2 run: state, last-step statistics and level-0 list at
    every 100th step (about a minute on one core)."""
    c2 = ics.symba_disk()
    o = O.OracleSystem(O.SYMBA, c2, 0.005, istep_out=100)
    out = []
    for _ in range(10):
        assert o.run(100) == 0
        out.append((o.state(), o.stats(), o.enc_list(0)[:2]))
    return c2, out


def test_c2_symba_disk_1e3_steps_bit_exact(c2_oracle):
    c2, ref = c2_oracle
    g = S.GpuSystem(S.SYMBA, c2, 0.005, istep_out=100)
    nsub = 0
    for k, (so, sto, (i1, i2)) in enumerate(ref):
        g.run(100)
        sg, stg = g.state(), g.stats()
        for key in ("pl_rh", "pl_vh", "pl_vb"):
            assert np.array_equal(so[key], sg[key]), (k, key)
        for key in ("nplpl_enc", "max_depth", "nsubsteps", "ncollision"):
            assert stg[key] == sto[key], (k, key, stg[key], sto[key])
        a, b = g.enc_list(0)
        assert np.array_equal(a, i1) and np.array_equal(b, i2), k
        nsub += stg["nsubsteps"]
    assert nsub > 0  # the recursion ran at the checkpoints


def test_c2_merge_mode_200_steps_bit_exact(c2_oracle):
    """C2 as specified, merge-only collisions: the same run with collisions
    merged inside the step (every level on the whole device, the host told
    of each substep's collision check).  The disk has no collision in its
    first 1000 steps, so the states equal the plain run's checkpoints."""
    c2, ref = c2_oracle
    g = S.GpuSystem(S.SYMBA, c2, 0.005, istep_out=100, merge=(ics.GU, None))
    for k, (so, sto, (i1, i2)) in enumerate(ref[:2]):
        g.run(100)
        sg, stg = g.state(), g.stats()
        for key in ("pl_rh", "pl_vh", "pl_vb"):
            assert np.array_equal(so[key], sg[key]), (k, key)
        for key in ("nplpl_enc", "max_depth", "nsubsteps"):
            assert stg[key] == sto[key], (k, key, stg[key], sto[key])
    assert g.nmerge == 0


def test_c2_fast_pairs_1e3_steps_within_1e_10(c2_oracle):
    c2, ref = c2_oracle
    g = S.GpuSystem(S.SYMBA, c2, 0.005, istep_out=100, pair_mode=S.PAIR_FAST)
    for k, (so, sto, (i1, i2)) in enumerate(ref):
        g.run(100)
        sg, stg = g.state(), g.stats()
        assert _rel(sg["pl_rh"], so["pl_rh"]) <= 1e-10, k
        assert stg["nplpl_enc"] == sto["nplpl_enc"], k
        a, b = g.enc_list(0)
        assert np.array_equal(a, i1) and np.array_equal(b, i2), k


def test_c2_fast_accelerations_within_1e_13():
    c2 = ics.symba_disk()
    f = S.GpuSystem(S.SYMBA, c2, 0.005, pair_mode=S.PAIR_FAST)
    e = S.GpuSystem(S.SYMBA, c2, 0.005)
    st = e.state()
    gm, R = c2.Gmass[st["pl_id"] - 1], c2.radius[st["pl_id"] - 1]
    pe, _ = e.accel()
    assert np.array_equal(pe, O.accel_plpl(st["pl_rh"], gm, radius=R))
    pf, _ = f.accel()
    assert _rel(pf, pe) <= 1e-13


# ------------------------------------------------------------------- C3 --

@pytest.fixture(scope="module")
def c3():
    return ics.symba_big()


def test_c3_level0_list_and_direct_sums(c3):
    g = S.GpuSystem(S.SYMBA, c3, 0.005, lclose=False)
    st = g.state()
    gm, R = c3.Gmass[st["pl_id"] - 1], c3.radius[st["pl_id"] - 1]
    # direct sums, exact mode: every row, and a sample of particles
    pe, te = g.accel()
    assert np.array_equal(pe, O.accel_plpl(st["pl_rh"], gm, radius=R, omp=True))
    idx = np.random.Generator(np.random.PCG64(3)).choice(c3.ntp, 65536, replace=False)
    assert np.array_equal(te[idx], O.accel_pltp(st["tp_rh"][idx], st["pl_rh"], gm, omp=True))
    # fast mode: within 1e-13 of the exact sums, and within 4e-15 of the
    # sum of the terms' magnitudes (the scale of the sums' own rounding)
    f = S.GpuSystem(S.SYMBA, c3, 0.005, lclose=False, pair_mode=S.PAIR_FAST)
    pf, tf = f.accel()
    del f
    assert _rel(pf, pe) <= 1e-13
    assert _rel(tf[idx], te[idx]) <= 1e-13
    sub = idx[:2048]
    scale = _abs_terms_tp(st["tp_rh"][sub], st["pl_rh"], gm)
    assert np.max(np.linalg.norm(tf[sub] - te[sub], axis=1) / scale) <= 4e-15
    # the level-0 pl-pl list of the first step: the check runs on the
    # state at the step's start with renc = rhill * RHSCALE
    renc = (st["rhill"] * 6.5) * 1.0
    i1, i2, _ = O.encounters_plpl(st["pl_rh"], st["pl_vb"], renc, 0.005, omp=True)
    g.step()
    a, b = g.enc_list(0)
    assert len(i1) > 10 ** 7
    assert np.array_equal(a, i1) and np.array_equal(b, i2)


def test_c3_drift_of_all_bodies(c3):
    r = np.concatenate([c3.pl_rh, c3.tp_rh])
    v = np.concatenate([c3.pl_vh, c3.tp_vh])
    mask = np.ones(len(r), np.int32)
    xg, vg, fg = S.swiftest_drift_all(c3.GMcb, r, v, len(r), 0.005, mask)
    xo, vo, fo = O.drift(c3.GMcb, r, v, 0.005, omp=True)
    assert np.array_equal(fg, fo)
    same = np.all(xg == xo, axis=1) & np.all(vg == vo, axis=1)
    # bit-identical except the drifts whose first guess calls libm (sin in
    # the Kepler guess), which agree to 1e-12
    assert same.mean() >= 0.99
    assert _rel(xg, xo) <= 1e-12 and _rel(vg, vo) <= 1e-12


# ------------------------------------------------------------------- C4 --

def test_c4_ensemble_10_steps_sampled_systems_equal_their_oracle_runs():
    systems = ics.ensemble()
    g = S.GpuSystem(S.SYMBA, systems, 0.01, istep_out=0, lclose=False)
    g.autotune(2)
    g.use_graphs(True)
    g.run(10)
    sg = g.state()
    off_pl = np.concatenate([[0], np.cumsum([s.npl for s in systems])])
    off_tp = np.concatenate([[0], np.cumsum([s.ntp for s in systems])])
    idx = np.random.Generator(np.random.PCG64(41)).choice(len(systems), 64, replace=False)
    for k in idx:
        s = systems[int(k)]
        o = O.OracleSystem(O.SYMBA, s, 0.01, lclose=False)
        assert o.run(10) == 0
        so = o.state()
        a, b = off_pl[k], off_pl[k + 1]
        assert np.array_equal(sg["pl_rh"][a:b], so["pl_rh"]), k
        assert np.array_equal(sg["pl_vb"][a:b], so["pl_vb"]), k
        a, b = off_tp[k], off_tp[k + 1]
        assert np.array_equal(sg["tp_rh"][a:b], so["tp_rh"]), k


# ------------------------------------------------------------ leaf, fast --

def test_leaf_operators_fast_mode():
    rng = np.random.Generator(np.random.PCG64(12))
    n, m = 300, 500
    r = rng.normal(0, 2, (n, 3))
    gm = 10 ** rng.uniform(-9, -4, n)
    R = 10 ** rng.uniform(-6, -3, n)
    rt = rng.normal(0, 2, (m, 3))
    try:
        S.set_pair_mode(S.PAIR_FAST)
        af = S.swiftest_kick_getacch_int_all_tri(n, n, r, gm, R, np.zeros((n, 3)))
        tf = S.swiftest_kick_getacch_int_all_tp(m, n, rt, r, gm, np.ones(m, np.int32), np.zeros((m, 3)))
    finally:
        S.set_pair_mode(S.PAIR_EXACT)
    ae = O.accel_plpl(r, gm, radius=R)
    te = O.accel_pltp(rt, r, gm)
    assert not np.array_equal(af, ae)   # the fast factor is in use
    assert _rel(af, ae) <= 1e-13 and _rel(tf, te) <= 1e-13
    # and the exact mode is back
    assert np.array_equal(S.swiftest_kick_getacch_int_all_tri(n, n, r, gm, R, np.zeros((n, 3))), ae)


# ------------------------------------- exact massive-body kernel variants --

@pytest.mark.parametrize("kev", [2, 4, 8])
def test_split_row_kernel_bit_exact(kev, monkeypatch):
    """The exact massive-body kick with its terms evaluated by several
    wavefronts and folded in order gives the one-lane-per-body sums bit for
    bit: accelerations of the 2000-body disk and of a crowded disk with
    encounter pairs removed, then 20 steps of a crowded disk against the
    oracle."""
    from helpers import compact_disk
    monkeypatch.setenv("SWG_KICK_PL_EVAL", str(kev))
    c2 = ics.symba_disk()
    g = S.GpuSystem(S.SYMBA, c2, 0.005)
    st = g.state()
    gm, R = c2.Gmass[st["pl_id"] - 1], c2.radius[st["pl_id"] - 1]
    pe, _ = g.accel()
    assert np.array_equal(pe, O.accel_plpl(st["pl_rh"], gm, radius=R))
    d = compact_disk(seed=32, npl=300, ntp=40)
    o = O.OracleSystem(O.SYMBA, d, 0.005)
    g = S.GpuSystem(S.SYMBA, d, 0.005)
    for _ in range(4):
        assert o.run(5) == 0
        g.run(5)
        so, sg = o.state(), g.state()
        for key in ("pl_rh", "pl_vb", "tp_rh", "tp_vb"):
            assert np.array_equal(so[key], sg[key]), key
        assert g.stats()["nplpl_enc"] == o.stats()["nplpl_enc"] > 0


# ------------------------------------------------ semi-interacting bodies --

@pytest.mark.parametrize("q,mode", [(0.5, "exact"), (0.8, "exact"), (0.8, "fast")])
def test_semi_interacting_bodies_match_the_oracle(q, mode):
    """GMTINY below the median (rows over the fully interacting bodies) and
    at the 80th percentile (the triangular branch): 30 SyMBA steps of a
    crowded disk with test particles against the oracle, bit for bit in the
    exact pair mode (1e-10 with the fast factor), the same encounter lists,
    and no pair of two semi-interacting bodies listed."""
    from helpers import compact_disk
    d = compact_disk(seed=38, npl=240, ntp=30, mlo=1e-9, mhi=1e-5)
    gmt = float(np.quantile(d.Gmass, q))
    pm = S.PAIR_FAST if mode == "fast" else S.PAIR_EXACT
    o = O.OracleSystem(O.SYMBA, d, 0.005, gmtiny=gmt)
    g = S.GpuSystem(S.SYMBA, d, 0.005, gmtiny=gmt, pair_mode=pm)
    nplm = int((d.Gmass >= gmt).sum())
    for k in range(6):
        assert o.run(5) == 0
        g.run(5)
        so, sg = o.state(), g.state()
        if mode == "exact":
            for key in ("pl_rh", "pl_vb", "tp_rh", "tp_vb"):
                assert np.array_equal(so[key], sg[key]), (k, key)
        else:
            assert _rel(sg["pl_rh"], so["pl_rh"]) <= 1e-10
        a, b = g.enc_list(0)
        i1, i2, _ = o.enc_list(0)
        assert np.array_equal(a, i1) and np.array_equal(b, i2), k
        assert np.all(a <= nplm)
    assert g.stats()["nplpl_enc"] > 0
