"""GPU parity of SyMBA steps whose collisions are resolved by merging inside
the step (swg_system_set_merge): the device stops at the substep whose
collision check flagged a pair, the host merges and rearranges the bodies,
and the step continues.  The oracle does the same in its recursion
(collision_resolve_plpl, collision_resolve.f90:521-662, with the MERGE model
and swiftest_util_rearray_pl, swiftest_util.f90:1638-1811).

Bar: bit-identical states, body ids, masses, radii, recursion counts and
level-0 lists after every block of steps.  The crowded disks below merge at
the deepest levels (where the rearray's encounter check over the full step
finds nothing and the old list is kept) and, with radii inflated 30x and
100x, at levels 2-6 (where it finds pairs, and the old entries' state is
carried over position by position and the list is cut to its old length).
"""
import numpy as np
import pytest

import pyoracle as O
import swiftest_amd as S
from helpers import compact_disk, ulp_diff
from swiftest_amd import ics

pytestmark = pytest.mark.gpu


def _compare(o, g, label):
    so, sg = o.state(), g.state()
    bo, bg = o.bodies(), g.bodies()
    assert so["pl_rh"].shape == sg["pl_rh"].shape, (label, so["pl_rh"].shape, sg["pl_rh"].shape)
    for k in ("id", "mass", "Gmass", "radius"):
        assert np.array_equal(bo[k], bg[k]), (label, k)
    assert np.array_equal(so["pl_id"], sg["pl_id"]), label
    for k in ("pl_rh", "pl_vh", "pl_vb", "pl_status", "rhill"):
        assert np.array_equal(so[k], sg[k]), f"{label} {k} ulp={ulp_diff(so[k], sg[k]) if so[k].dtype == np.float64 else ''}"
    st_o, st_g = o.stats(), g.stats()
    for k in ("nplpl_enc", "max_depth", "nsubsteps", "ncollision"):
        assert st_o[k] == st_g[k], (label, k, st_o, st_g)
    i1, i2, _ = o.enc_list(0)
    a, b = g.enc_list(0)
    assert np.array_equal(i1, a) and np.array_equal(i2, b), label
    assert o.nmerge == g.nmerge, (label, o.nmerge, g.nmerge)


@pytest.mark.parametrize("seed,scale,min_merges", [(11, 1.0, 1), (11, 30.0, 3), (13, 100.0, 9)])
def test_merges_bit_exact(seed, scale, min_merges):
    sysic = compact_disk(seed=seed, npl=160, ntp=0)
    sysic.radius[:] = sysic.radius * scale
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, istep_out=10, merge=(ics.GU, None))
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, istep_out=10, merge=(ics.GU, None))
    nblk = 6 if scale == 1.0 else 4
    for blk in range(nblk):
        print(f"merge seed={seed} scale={scale} block {blk}", flush=True)
        assert o.run(10) == 0
        g.run(10)
        _compare(o, g, f"seed {seed} scale {scale} block {blk}")
    assert g.nmerge >= min_merges
    b = g.bodies()
    assert len(b["id"]) == 160 - g.nmerge
    assert set(b["id"][b["id"] > 160]) <= set(range(161, 161 + g.nmerge))


@pytest.mark.parametrize("seed,scale", [(11, 30.0), (13, 100.0)])
def test_merges_with_semi_interacting_bodies(seed, scale):
    """Merges in a disk whose lighter half is semi-interacting (GMTINY at the
    median Gmass): a merged body is fully interacting or not by its own
    Gmass, the count of fully interacting bodies follows the rearray, and
    every block equals the oracle's bit for bit."""
    sysic = compact_disk(seed=seed, npl=160, ntp=0)
    sysic.radius[:] = sysic.radius * scale
    gt = float(np.median(sysic.Gmass))
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, istep_out=10, merge=(ics.GU, None), gmtiny=gt)
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, istep_out=10, merge=(ics.GU, None), gmtiny=gt)
    for blk in range(4):
        print(f"merge gmtiny seed={seed} scale={scale} block {blk}", flush=True)
        assert o.run(10) == 0
        g.run(10)
        _compare(o, g, f"gmtiny seed {seed} scale {scale} block {blk}")
    assert g.nmerge >= 3
    b = g.bodies()
    assert 0 < (b["Gmass"] < gt).sum() < len(b["id"])


def test_merges_with_discard_checks_and_energy():
    """Merges between discard checks: the pericentre state of the checks
    (isperi, atp, peri per body) follows the bodies through the rearray, so
    the checks after each step give the oracle's counts and state; and the
    energy record after the merges, with the merged bodies' masses, equals
    the oracle's energy of the same state to 1e-12."""
    sysic = compact_disk(seed=13, npl=160, ntp=0)
    sysic.radius[:] = sysic.radius * 100.0
    kw = dict(rmin=0.3, rmax=5.0, qmin=0.5, qmin_alo=0.1, qmin_ahi=10.0, qmin_coord=0)
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, merge=(ics.GU, None))
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, merge=(ics.GU, None))
    for step in range(30):
        print(f"merge+discard step {step}", flush=True)
        assert o.step() == 0
        g.step()
        assert g.discard(**kw) == o.discard(**kw) == (0, 0), step
        _compare(o, g, f"step {step}")
    assert g.nmerge >= 5
    st, b = o.state(), o.bodies()
    cbm = sysic.GMcb / ics.GU
    got = g.energy(cbm, np.array([sysic.Rcb]))[0]
    ref, _ = O.energy(sysic.GMcb, cbm, sysic.Rcb, st["cb_vb"], st["pl_rh"], st["pl_vb"], b["Gmass"], b["mass"],
                      b["radius"], st["pl_status"], lclose=True)
    for f in (0, 2, 3, 4, 5, 15):
        assert got[f] == pytest.approx(ref[f], rel=1e-12, abs=0), f
    for a, c in ((6, 9), (12, 15)):
        assert np.linalg.norm(got[a:c] - ref[a:c]) <= 1e-12 * np.linalg.norm(ref[a:c]), a


def test_merges_in_an_ensemble():
    """Three crowded disks in one ensemble, each system's merges its own: a
    merge shrinks its system and moves the later systems' bodies down.
    Every system equals its own oracle merge-mode run, bit for bit."""
    systems = []
    for seed, scale in ((11, 30.0), (12, 1.0), (13, 100.0)):
        s = compact_disk(seed=seed, npl=160 if seed != 12 else 120, ntp=0)
        s.radius[:] = s.radius * scale
        systems.append(s)
    g = S.GpuSystem(S.SYMBA, systems, 0.005, istep_out=10, merge=(ics.GU, None))
    orc = [O.OracleSystem(O.SYMBA, s, 0.005, istep_out=10, merge=(ics.GU, None)) for s in systems]
    for blk in range(4):
        print(f"ensemble merges block {blk}", flush=True)
        g.run(10)
        for o in orc:
            assert o.run(10) == 0
        sg, bg = g.state(), g.bodies()
        off = np.concatenate([[0], np.cumsum(g.npl_s)])
        assert g.nmerge == sum(o.nmerge for o in orc)
        for k, o in enumerate(orc):
            sl = slice(off[k], off[k + 1])
            so, bo = o.state(), o.bodies()
            for key in ("id", "mass", "Gmass", "radius"):
                assert np.array_equal(bg[key][sl], bo[key]), (blk, k, key)
            for key in ("pl_rh", "pl_vh", "pl_vb", "rhill"):
                assert np.array_equal(sg[key][sl], so[key]), (blk, k, key, ulp_diff(sg[key][sl], so[key]))
            assert np.array_equal(sg["cb_vb"][k], so["cb_vb"]), (blk, k)
    assert g.nmerge >= 10


def test_merge_mode_refusals():
    """Outside its scope the merge mode is refused at set-up, with an error
    and no change to the system."""
    disk = compact_disk(seed=11, npl=40, ntp=0)
    with_tp = compact_disk(seed=11, npl=40, ntp=8)
    cases = [lambda: S.GpuSystem(S.HELIO, disk, 0.005, merge=(ics.GU, None)),
             lambda: S.GpuSystem(S.SYMBA, with_tp, 0.005, merge=(ics.GU, None)),
             lambda: S.GpuSystem(S.SYMBA, disk, 0.005, lclose=False, merge=(ics.GU, None)),
             lambda: S.GpuSystem(S.SYMBA, disk, 0.005, encounter_check=1, merge=(ics.GU, None)),
             lambda: S.GpuSystem(S.SYMBA, disk, 0.005, merge=(0.0, None))]
    for make in cases:
        with pytest.raises(S.SwiftestHipError):
            make()
    pool = S.GpuSystem.slots(S.SYMBA, [64], [8], 0.005)  # room for test particles
    assert pool.L.swg_system_set_merge(pool.h, ics.GU, None) != 0
    pool = S.GpuSystem.slots(S.SYMBA, [64], [0], 0.005)  # a slot pool takes no masses
    mass = np.ones(64)
    assert pool.L.swg_system_set_merge(pool.h, ics.GU, mass.ctypes.data) != 0


def test_merges_in_a_slot_pool():
    """Merges in a slot pool (continuous batching): a slot that merges keeps
    its range and pads what its system no longer fills, a slot loaded later
    runs its own system from that point, and a slot reloaded after merges
    gives its new system fresh ids and masses.  Each system equals the
    oracle's merge-mode run of it alone, bit for bit."""
    def disk(seed, npl, scale):
        s = compact_disk(seed=seed, npl=npl, ntp=0)
        s.radius[:] = s.radius * scale
        return s

    pool = S.GpuSystem.slots(S.SYMBA, [170, 130, 170], [0, 0, 0], 0.005, merge_gu=ics.GU)
    orc = {}
    done_merges = 0

    def load(k, s):
        nonlocal done_merges
        if k in orc:
            done_merges += orc[k].nmerge
        pool.load(k, s)
        orc[k] = O.OracleSystem(O.SYMBA, s, 0.005, merge=(ics.GU, None))

    load(0, disk(11, 160, 30.0))
    load(2, disk(13, 160, 100.0))
    for blk in range(4):
        if blk == 2:
            load(1, disk(12, 120, 1.0))     # a slot that was empty
            load(2, disk(14, 150, 100.0))   # a slot whose system merged
        print(f"slot-pool merges block {blk}", flush=True)
        pool.run(10)
        for o in orc.values():
            assert o.run(10) == 0
        for k, o in orc.items():
            sg, bg = pool.get_slot(k), pool.slot_bodies(k)
            so, bo = o.state(), o.bodies()
            assert pool.slot_counts(k)[0] == len(bo["id"]), (blk, k)
            for key in ("id", "mass", "Gmass", "radius"):
                assert np.array_equal(bg[key], bo[key]), (blk, k, key)
            for key in ("pl_rh", "pl_vh", "pl_vb", "pl_status", "pl_id"):
                assert np.array_equal(sg[key], so[key]), (blk, k, key)
        assert pool.nmerge == done_merges + sum(o.nmerge for o in orc.values())
    assert orc[2].nmerge >= 1 and done_merges >= 3
