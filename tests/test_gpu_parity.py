"""GPU parity: the HIP path against the strict CPU restatement on identical
seeded inputs.

Bar: every kernel whose arithmetic is +, -, *, /, sqrt and fma is required to
be bit-identical to the oracle (accelerations, elliptic Danby drift,
encounter lists, whole trajectories).  Two drift branches call libm
transcendental functions (sin in the Kepler guess when dt/r0 > 0.4, x**(1/3)
in the hyperbolic cubic guess); there the device and the host libm may
differ by an ulp in the starting guess, and the converged result is checked
to 1e-12 relative.
"""
import numpy as np
import pytest

import pyoracle as O
import swiftest_amd as S
from helpers import compact_disk, sparse_system, ulp_diff
from swiftest_amd import ics

pytestmark = pytest.mark.gpu


def _bodies(seed, n, spread=2.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.normal(0, spread, (n, 3))
    Gm = ics.GU * np.exp(rng.uniform(np.log(1e-9), np.log(1e-3), n))
    R = np.exp(rng.uniform(np.log(1e-5), np.log(1e-2), n))
    return r, Gm, R


@pytest.mark.parametrize("n,nplm", [(257, 257), (130, 100), (90, 20)])
@pytest.mark.parametrize("with_radius", [True, False])
def test_kick_tri_bit_exact(n, nplm, with_radius):
    print(f"kick_tri n={n} nplm={nplm} radius={with_radius}", flush=True)
    r, Gm, R = _bodies(1 + n, n)
    r[5] = r[4] + 1e-3  # a planted overlapping pair (skipped when radii are used)
    acc0 = np.random.Generator(np.random.PCG64(7)).normal(0, 1e-3, (n, 3))
    rad = R if with_radius else None
    ref = O.accel_plpl(r, Gm, rad, acc0, nplm=nplm)
    got = S.swiftest_kick_getacch_int_all_tri(n, nplm, r, Gm, rad, acc0)
    assert np.array_equal(ref, got), ulp_diff(ref, got)


@pytest.mark.parametrize("with_radius", [True, False])
def test_kick_flat_bit_exact(with_radius):
    n = 200
    r, Gm, R = _bodies(3, n)
    rng = np.random.Generator(np.random.PCG64(9))
    k = []
    for _ in range(1500):
        i, j = sorted(rng.choice(n, 2, replace=False))
        k.append((i + 1, j + 1))
    k = np.array(k, np.int32)
    acc0 = rng.normal(0, 1e-3, (n, 3))
    rad = R if with_radius else None
    ref = O.accel_pairs(k, r, Gm, rad, acc0)
    got = S.swiftest_kick_getacch_int_all_flat(n, k, r, Gm, rad, acc0)
    assert np.array_equal(ref, got), ulp_diff(ref, got)


@pytest.mark.parametrize("npl,ntp", [(300, 5000), (1, 64), (513, 257)])
def test_kick_tp_bit_exact(npl, ntp):
    print(f"kick_tp npl={npl} ntp={ntp}", flush=True)
    rpl, Gm, _ = _bodies(5, npl)
    rtp, _, _ = _bodies(6, ntp, spread=3.0)
    rng = np.random.Generator(np.random.PCG64(8))
    lmask = (rng.uniform(size=ntp) > 0.1).astype(np.int32)
    acc0 = rng.normal(0, 1e-3, (ntp, 3))
    ref = O.accel_pltp(rtp, rpl, Gm, lmask, acc0)
    got = S.swiftest_kick_getacch_int_all_tp(ntp, npl, rtp, rpl, Gm, lmask, acc0)
    assert np.array_equal(ref, got), ulp_diff(ref, got)


def test_pair_factor_outside_the_fast_range_is_still_ieee():
    """The pair factor Gm / (r2 sqrt(r2)) takes a shortened sequence only for
    separations and masses well inside the double range; separations from
    1e-150 to 1e150 AU and masses of 0, 1e-170 and 1e170 (both sides of that
    range) must give the IEEE result bit for bit."""
    rng = np.random.Generator(np.random.PCG64(31))
    npl, ntp = 70, 600
    rpl = rng.normal(0, 1.0, (npl, 3))
    Gm = ics.GU * np.exp(rng.uniform(np.log(1e-9), np.log(1e-3), npl))
    Gm[3], Gm[10], Gm[40] = 0.0, 1e-170, 1e170
    scale = 10.0 ** rng.uniform(-150, 150, ntp)
    rtp = rpl[rng.integers(0, npl, ntp)] + rng.normal(0, 1.0, (ntp, 3)) * scale[:, None]
    lmask = np.ones(ntp, np.int32)
    acc0 = np.zeros((ntp, 3))
    with np.errstate(all="ignore"):
        ref = O.accel_pltp(rtp, rpl, Gm, lmask, acc0)
    got = S.swiftest_kick_getacch_int_all_tp(ntp, npl, rtp, rpl, Gm, lmask, acc0)
    assert np.array_equal(ref, got, equal_nan=True), ulp_diff(ref, got)
    rtri = np.concatenate([rpl[:30], rpl[:30] + rng.normal(0, 1.0, (30, 3)) * (10.0 ** rng.uniform(-150, 150, 30))[:, None]])
    gtri = np.concatenate([Gm[:30], Gm[30:60]])
    with np.errstate(all="ignore"):
        ref = O.accel_plpl(rtri, gtri)
    got = S.swiftest_kick_getacch_int_all_tri(60, 60, rtri, gtri, None, np.zeros((60, 3)))
    assert np.array_equal(ref, got, equal_nan=True), ulp_diff(ref, got)


def _orbits(seed, n, amin, amax, emax, hyper=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.uniform(amin, amax, n)
    e = rng.uniform(0, emax, n)
    inc = rng.uniform(0, 0.3, n)
    angs = [rng.uniform(0, 2 * np.pi, n) for _ in range(3)]
    r, v = ics.el2xv(ics.GMSUN, a, e, inc, *angs)
    if hyper:
        v = v * rng.uniform(1.5, 3.0, (n, 1))
    return r, v


@pytest.mark.parametrize("dt", [0.005, 0.1, 2.0])
def test_drift_elliptic(dt):
    print(f"drift dt={dt}", flush=True)
    r, v = _orbits(12, 4000, 0.3, 30.0, 0.5)
    mu = np.full(len(r), ics.GMSUN)
    lmask = np.ones(len(r), np.int32)
    rx, vx, fx = O.drift(mu, r, v, dt, lmask)
    gx, gv, gf = S.swiftest_drift_all(mu, r, v, len(r), dt, lmask)
    assert np.array_equal(fx, gf)
    dtr0 = dt / np.linalg.norm(r, axis=1)
    exact = dtr0 <= 0.4   # the series guess; no libm call on this path
    assert np.array_equal(rx[exact], gx[exact]), ulp_diff(rx[exact], gx[exact])
    assert np.array_equal(vx[exact], gv[exact]), ulp_diff(vx[exact], gv[exact])
    if (~exact).any():
        np.testing.assert_allclose(gx[~exact], rx[~exact], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(gv[~exact], vx[~exact], rtol=1e-12, atol=1e-14)
        same = np.all(gx[~exact] == rx[~exact], axis=1) & np.all(gv[~exact] == vx[~exact], axis=1)
        print(f"sin-guess drifts: {(~exact).sum()}, bit-identical {same.mean():.5f}", flush=True)
        # the device's and libm's sin may differ in the guess's last bit;
        # the converged iteration almost always removes the difference
        assert same.mean() >= 0.99, same.mean()


@pytest.mark.parametrize("dt", [0.05, 1e-6, 2.8e-8])
def test_drift_hyperbolic(dt):
    """Unbound orbits take the cubic first guess, whose two cube roots the
    device rounds correctly (pow_third) while the oracle calls libm's pow,
    which misrounds about 0.08 % of them: all drifts agree to 1e-12, and
    nearly all of them bit for bit."""
    print(f"hyperbolic drift dt={dt}", flush=True)
    r, v = _orbits(13, 2000, 0.5, 5.0, 0.3, hyper=True)
    mu = np.full(len(r), ics.GMSUN)
    lmask = np.ones(len(r), np.int32)
    rx, vx, fx = O.drift(mu, r, v, dt, lmask)
    gx, gv, gf = S.swiftest_drift_all(mu, r, v, len(r), dt, lmask)
    assert np.array_equal(fx, gf)
    np.testing.assert_allclose(gx, rx, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(gv, vx, rtol=1e-12, atol=1e-14)
    same = np.all(gx == rx, axis=1) & np.all(gv == vx, axis=1)
    print(f"bit-identical drifts: {same.mean():.5f}", flush=True)
    assert same.mean() >= 0.99, same.mean()


@pytest.mark.parametrize("n", [1, 63, 64, 65, 700])
def test_encounter_plpl_exact(n):
    print(f"encounter plpl n={n}", flush=True)
    rng = np.random.Generator(np.random.PCG64(20 + n))
    r = rng.uniform(-1, 1, (n, 3))
    v = rng.normal(0, 1.0, (n, 3))
    renc = rng.uniform(0.01, 0.08, n)
    ref = O.encounters_plpl(r, v, renc, 0.05)
    got = S.encounter_check_all_plpl(n, r, v, renc, 0.05)
    assert len(ref[0]) == len(got[0])
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("npl,ntp", [(40, 1000), (130, 65), (1, 1)])
def test_encounter_pltp_exact(npl, ntp):
    print(f"encounter pltp npl={npl} ntp={ntp}", flush=True)
    rng = np.random.Generator(np.random.PCG64(40 + npl))
    rp = rng.uniform(-1, 1, (npl, 3))
    vp = rng.normal(0, 1.0, (npl, 3))
    rt = rng.uniform(-1, 1, (ntp, 3))
    vt = rng.normal(0, 1.0, (ntp, 3))
    renc = rng.uniform(0.02, 0.2, npl)
    ref = O.encounters_pltp(rp, vp, rt, vt, renc, 0.05)
    got = S.encounter_check_all_pltp(npl, ntp, rp, vp, rt, vt, renc, 0.05)
    assert len(ref[0]) == len(got[0])
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def _compare_systems(o, g, label, whm=False):
    so, sg = o.state(), g.state()
    assert np.array_equal(so["pl_id"], sg["pl_id"]), label
    keys = ("pl_rh", "pl_vh", "tp_rh", "tp_vh", "pl_status", "tp_status") + (() if whm else ("pl_vb",))
    for k in keys:
        assert np.array_equal(so[k], sg[k]), f"{label} {k} ulp={ulp_diff(so[k], sg[k]) if so[k].dtype == np.float64 else ''}"


def _compare_collision(o, g, label):
    """Both sides stopped for a collision at the end of the same step: the
    same level-0 lists with the same pairs flagged COLLIDED (at least one),
    the same recursion counts and the same state, bit for bit."""
    so, sg = o.stats(), g.stats()
    for k in ("nplpl_enc", "npltp_enc", "max_depth", "nsubsteps"):
        assert so[k] == sg[k], (label, k, so, sg)
    ncol = 0
    for which in (0, 1):
        i1, i2, _ = o.enc_list(which)
        a, b = g.enc_list(which)
        assert np.array_equal(i1, a) and np.array_equal(i2, b), (label, which)
        st_o, st_g = o.enc_status(which), g.enc_status(which)
        assert np.array_equal(st_o, st_g), (label, which, np.flatnonzero(st_o != st_g)[:10])
        ncol += int((st_o == -14).sum())
    assert ncol > 0, label
    _compare_systems(o, g, label)
    return ncol


@pytest.mark.parametrize("short_circuit,ntp", [(True, 200), (False, 40)])
def test_symba_crowded_disk_bit_exact(short_circuit, ntp):
    sysic = compact_disk(npl=160, ntp=ntp)
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, short_circuit=short_circuit, istep_out=10, collision_finish_step=True)
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, short_circuit=short_circuit, istep_out=10)
    depth = -1
    for blk in range(4):
        print(f"symba crowded disk sc={short_circuit} block {blk}", flush=True)
        rc = o.run(10)
        try:
            g.run(10)
            grc = 0
        except S.SwiftestHipError as e:
            grc = e.code
        # a collision (merge is host work) ends both runs at the end of the
        # same step, with the same pairs flagged and the same state
        assert (rc == -1) == (grc == -3), (rc, grc)
        if rc:
            _compare_collision(o, g, f"collision in block {blk}")
            break
        so, sg = o.stats(), g.stats()
        assert so["nplpl_enc"] == sg["nplpl_enc"] and so["npltp_enc"] == sg["npltp_enc"]
        assert so["max_depth"] == sg["max_depth"] and so["nsubsteps"] == sg["nsubsteps"], (so, sg)
        depth = max(depth, so["max_depth"])
        _compare_systems(o, g, f"block {blk}")
    assert depth >= 1


@pytest.mark.parametrize("integ", [S.HELIO, S.SYMBA, S.WHM])
def test_sparse_system_bit_exact(integ):
    print(f"sparse system integ={integ}", flush=True)
    sysic = sparse_system()
    o = O.OracleSystem(integ, sysic, 0.01, istep_out=25)
    g = S.GpuSystem(integ, sysic, 0.01, istep_out=25)
    assert o.run(100) == 0
    g.run(100)
    _compare_systems(o, g, f"integrator {integ}", whm=integ == S.WHM)


def test_level0_lists_match():
    sysic = compact_disk(seed=5, npl=300, ntp=500, alo=0.9, ahi=1.1)
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, short_circuit=False)
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, short_circuit=False)
    for _ in range(3):
        o.step()
        g.step()
        for which in (0, 1):
            a1, a2, _ = o.enc_list(which)
            b1, b2 = g.enc_list(which)
            assert np.array_equal(a1, b1) and np.array_equal(a2, b2)
    _compare_systems(o, g, "lists")


@pytest.mark.parametrize("subtree_max,lds,wave", [("-1", "1", "1"), ("60", "1", "1"), ("60", "1", "0"),
                                                   ("60", "0", "1"), ("1000000000", "1", "1"),
                                                   ("1000000000", "1", "0"), ("1000000000", "0", "1")])
def test_recursion_modes_match_oracle(monkeypatch, subtree_max, lds, wave):
    """Every level on the whole device (-1), a hand-off to the per-system
    persistent workgroups part-way down (60) and right at level 1 (1e9), with
    the LDS-resident workgroup kernel on or off and its one-wavefront mode for
    small deep levels on or off, all reproduce the oracle bit for bit, with
    the same substep count."""
    print(f"recursion mode SWG_SUBTREE_MAX={subtree_max} SWG_SUBTREE_LDS={lds} SWG_SUBTREE_WAVE={wave}", flush=True)
    sysic = compact_disk(seed=12, npl=200, ntp=150, mhi=2e-5)
    monkeypatch.setenv("SWG_SUBTREE_MAX", subtree_max)
    monkeypatch.setenv("SWG_SUBTREE_LDS", lds)
    monkeypatch.setenv("SWG_SUBTREE_WAVE", wave)
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, short_circuit=False, lclose=False, istep_out=5)
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, short_circuit=False, lclose=False, istep_out=5)
    for blk in range(3):
        assert o.run(5) == 0
        g.run(5)
        so, sg = o.stats(), g.stats()
        assert (so["max_depth"], so["nsubsteps"]) == (sg["max_depth"], sg["nsubsteps"]), (blk, so, sg)
        _compare_systems(o, g, f"block {blk}")
    assert sg["max_depth"] >= 2


@pytest.mark.parametrize("npl,mhi,ntp", [(400, 1e-4, 100), (600, 3e-5, 200)])
def test_deep_recursion_with_tp_bit_exact(npl, mhi, ntp):
    """Heavier crowded disks with test particles in encounters (no short
    circuit) reach depth 8-9: the default scheduling (whole-device levels,
    hand-off, the LDS-resident workgroup and its one-wavefront mode, the kept
    closing-kick sums) reproduces the oracle bit for bit, substep for
    substep, over 20 steps."""
    sysic = compact_disk(seed=31, npl=npl, ntp=ntp, mhi=mhi)
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, short_circuit=False, lclose=False, istep_out=10)
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, short_circuit=False, lclose=False, istep_out=10)
    for blk in range(2):
        print(f"deep recursion npl={npl} ntp={ntp} block {blk}", flush=True)
        assert o.run(10) == 0
        g.run(10)
        so, sg = o.stats(), g.stats()
        assert (so["max_depth"], so["nsubsteps"], so["npltp_enc"]) == (sg["max_depth"], sg["nsubsteps"],
                                                                       sg["npltp_enc"]), (blk, so, sg)
        _compare_systems(o, g, f"block {blk}")
    assert sg["max_depth"] >= 6


def _sas_case(kind):
    if kind == "disk":
        return [compact_disk(seed=13, npl=300, ntp=0)], 0.005
    rng = np.random.Generator(np.random.PCG64(17))
    if kind == "clustered":
        # a few tight clumps: many bodies share x extents
        base = compact_disk(seed=14, npl=240, ntp=0)
        centers = base.pl_rh[rng.integers(0, 6, base.npl)]
        base.pl_rh[:] = centers + rng.normal(0, 2e-3, (base.npl, 3))
        return [base], 0.005
    if kind == "fast":
        # large velocities and a long step: the motion over dt decides the pairs
        base = compact_disk(seed=15, npl=200, ntp=0)
        base.pl_vh[:] = base.pl_vh * rng.uniform(0.5, 3.0, (base.npl, 1)) + rng.normal(0, 3.0, (base.npl, 3))
        return [base], 0.05
    # an ensemble of unequal systems (sorted and swept per system)
    return [compact_disk(seed=20 + k, npl=n, ntp=0) for k, n in enumerate((40, 130, 7, 90))], 0.005


@pytest.mark.parametrize("kind", ["disk", "clustered", "fast", "ensemble"])
def test_sortsweep_lists_equal_triangular(kind):
    """The sort-and-sweep broad phase is conservative: after the same narrow
    phase its level-0 pl-pl lists equal the triangular (all-pairs) lists,
    step after step, and so do the trajectories."""
    print(f"sortsweep {kind}", flush=True)
    systems, dt = _sas_case(kind)
    tri = S.GpuSystem(S.SYMBA, systems, dt, lclose=False, encounter_check=0)
    sas = S.GpuSystem(S.SYMBA, systems, dt, lclose=False, encounter_check=1)
    total = 0
    for step in range(3):
        tri.step()
        sas.step()
        a1, a2 = tri.enc_list(0)
        b1, b2 = sas.enc_list(0)
        total += len(a1)
        assert np.array_equal(a1, b1) and np.array_equal(a2, b2), (kind, step)
        assert tri.stats()["nsubsteps"] == sas.stats()["nsubsteps"]
    assert total > 0
    a, b = tri.state(), sas.state()
    for k in ("pl_rh", "pl_vh", "pl_vb"):
        assert np.array_equal(a[k], b[k]), k


def test_sortsweep_symba_matches_oracle():
    sysic = compact_disk(npl=160, ntp=40)
    o = O.OracleSystem(O.SYMBA, sysic, 0.005, short_circuit=False, istep_out=10, collision_finish_step=True)
    g = S.GpuSystem(S.SYMBA, sysic, 0.005, short_circuit=False, istep_out=10, encounter_check=1)
    for blk in range(2):
        rc = o.run(10)
        try:
            g.run(10)
            grc = 0
        except S.SwiftestHipError as e:
            grc = e.code
        assert (rc == -1) == (grc == -3), (rc, grc)
        if rc:
            _compare_collision(o, g, f"sortsweep collision in block {blk}")
            break
        _compare_systems(o, g, f"sortsweep block {blk}")


@pytest.mark.parametrize("integ", ["symba", "helio", "ensemble"])
def test_graph_replays_equal_direct_launches(integ):
    """The step halves replayed from captured HIP graphs give the same state
    as direct launches, checked after every replay, with host copies and a
    freeze / unfreeze (new tiles, so new graphs) in between."""
    print(f"graphs {integ}", flush=True)
    if integ == "ensemble":
        systems = [compact_disk(seed=30 + k, npl=n, ntp=t) for k, (n, t) in enumerate(((50, 20), (90, 0), (12, 40)))]
        kind = S.SYMBA
    else:
        systems = [compact_disk(seed=31, npl=120, ntp=30)]
        kind = S.SYMBA if integ == "symba" else S.HELIO
    eager = S.GpuSystem(kind, systems, 0.005, lclose=False, istep_out=4)
    graph = S.GpuSystem(kind, systems, 0.005, lclose=False, istep_out=4)
    graph.use_graphs(True)
    for step in range(12):
        if integ == "ensemble" and step in (5, 8):
            f = np.array([0, 1, 0] if step == 5 else [0, 0, 0], np.int32)
            eager.set_frozen(f)
            graph.set_frozen(f)
        eager.run(1)
        graph.run(1)
        a, b = eager.state(), graph.state()
        for k in ("pl_rh", "pl_vh", "pl_vb", "tp_rh", "tp_vh", "tp_vb"):
            assert np.array_equal(a[k], b[k]), (step, k)


def test_ensemble_equals_single_runs():
    systems = [compact_disk(seed=s, npl=n, ntp=t, mhi=2e-6, alo=0.9, ahi=1.1)
               for s, n, t in ((1, 40, 30), (2, 70, 0), (3, 5, 90), (6, 120, 10))]
    ens = S.GpuSystem(S.SYMBA, systems, 0.005, istep_out=5)
    ens.run(20)
    se = ens.state()
    po = to = 0
    for s in systems:
        o = O.OracleSystem(O.SYMBA, s, 0.005, istep_out=5)
        assert o.run(20) == 0
        so = o.state()
        for k, n, off in (("pl_rh", s.npl, po), ("pl_vh", s.npl, po), ("tp_rh", s.ntp, to), ("tp_vh", s.ntp, to)):
            assert np.array_equal(so[k], se[k][off:off + n]), k
        po += s.npl
        to += s.ntp


# ---- energy and angular momentum (swg_system_energy) ----------------------
# Tolerance: the device sums in tree order and takes |r| as sqrt of the dot
# product, the oracle follows the reference's serial sums and gfortran's
# scaled norm2, so scalar fields agree to 1e-12 relative and the angular
# momentum vector to 1e-12 of its norm.

@pytest.mark.parametrize("integ", [S.SYMBA, S.HELIO, S.WHM])
def test_energy_matches_oracle(integ):
    print(f"energy integ={integ}", flush=True)
    systems = [compact_disk(seed=31, npl=600, mhi=2e-6, alo=0.8, ahi=1.2), compact_disk(seed=32, npl=7),
               sparse_system(seed=33, npl=9, ntp=20), compact_disk(seed=34, npl=2300, mhi=1e-6, alo=0.5, ahi=3.0)]
    g = S.GpuSystem(integ, systems, 0.002, lclose=True)
    g.run(3)
    st = g.state()
    rng = np.random.Generator(np.random.PCG64(35))
    cbm = np.array([s.GMcb / ics.GU for s in systems]) * rng.uniform(0.9, 1.1, len(systems))
    cbR = np.array([s.Rcb for s in systems])
    gm = np.concatenate([s.Gmass[st["pl_id"][o:o + s.npl] - 1] for s, o in
                         zip(systems, np.cumsum([0] + [s.npl for s in systems[:-1]]))])
    mass = gm / ics.GU * rng.uniform(0.99, 1.01, gm.shape[0])
    got = g.energy(cbm, cbR, mass)
    assert np.array_equal(got, g.energy(cbm, cbR, mass))  # same bits every call
    o = 0
    for k, s in enumerate(systems):
        sl = slice(o, o + s.npl)
        if integ == S.WHM:
            vb, cbv = O.vh2vb(s.GMcb, gm[sl], st["pl_vh"][sl])
        else:
            vb, cbv = st["pl_vb"][sl], st["cb_vb"][k]
        ref, _ = O.energy(s.GMcb, cbm[k], cbR[k], cbv, st["pl_rh"][sl], vb, gm[sl], mass[sl], s.radius[st["pl_id"][sl] - 1],
                          st["pl_status"][sl], lclose=True)
        for f in (0, 2, 3, 4, 5, 15):
            assert got[k, f] == pytest.approx(ref[f], rel=1e-12, abs=0), (k, f)
        for a, b in ((6, 9), (12, 15)):
            assert np.linalg.norm(got[k, a:b] - ref[a:b]) <= 1e-12 * np.linalg.norm(ref[a:b]), (k, a)
        assert got[k, 1] == 0 and np.all(got[k, 9:12] == 0)
        o += s.npl


# ---- 10^3 steps: positions and the energy-error trace ---------------------
# Tolerances: positions and velocities within 1e-10 of the oracle's (relative
# to the largest component), after every block of 100 steps up to step 1000;
# encounter counts, depth and substeps equal.  The relative energy error
# (E - E0) / |E0| from the device's energy reduction on the device's
# trajectory against the oracle's on its own: within 2e-12 at every block
# (each energy agrees to 1e-12 relative, so the two traces agree to the order
# of the error itself or better).

def _energy_of(st, sysic, integ, on_device, g=None):
    gm = sysic.Gmass[st["pl_id"] - 1]
    mass = gm / ics.GU
    cbm = sysic.GMcb / ics.GU
    if on_device:
        return g.energy(np.array([cbm]), np.array([sysic.Rcb]), mass)[0, 5]
    if integ == S.WHM:
        vb, cbv = O.vh2vb(sysic.GMcb, gm, st["pl_vh"])
    else:
        vb, cbv = st["pl_vb"], st["cb_vb"]
    rec, _ = O.energy(sysic.GMcb, cbm, sysic.Rcb, cbv, st["pl_rh"], vb, gm, mass,
                      sysic.radius[st["pl_id"] - 1], st["pl_status"])
    return rec[5]


@pytest.mark.parametrize("integ,kind", [(S.SYMBA, "crowded"), (S.SYMBA, "crowded-tp"), (S.HELIO, "sparse"),
                                        (S.WHM, "sparse")])
def test_thousand_steps_positions_and_energy_trace(integ, kind):
    if kind == "crowded":
        sysic, dt = compact_disk(seed=41, npl=120, ntp=50), 0.005
    elif kind == "crowded-tp":
        sysic, dt = compact_disk(seed=44, npl=100, ntp=200), 0.005
    else:
        sysic, dt = sparse_system(seed=42, npl=12, ntp=300), 0.01
    o = O.OracleSystem(integ, sysic, dt, lclose=False, istep_out=100)
    g = S.GpuSystem(integ, sysic, dt, lclose=False, istep_out=100)
    e0o = _energy_of(o.state(), sysic, integ, False)
    e0g = _energy_of(g.state(), sysic, integ, True, g)
    assert e0g == pytest.approx(e0o, rel=1e-12, abs=0)
    depth = -1
    for blk in range(10):
        print(f"1000 steps integ={integ} {kind}: steps {100 * blk}..{100 * blk + 99}", flush=True)
        assert o.run(100) == 0
        g.run(100)
        so, sg = o.state(), g.state()
        if integ == S.SYMBA:  # encounter statistics of the recursion
            ss, gs = o.stats(), g.stats()
            assert (ss["nplpl_enc"], ss["npltp_enc"], ss["max_depth"], ss["nsubsteps"]) == \
                (gs["nplpl_enc"], gs["npltp_enc"], gs["max_depth"], gs["nsubsteps"]), (blk, ss, gs)
            depth = max(depth, ss["max_depth"])
        assert np.array_equal(so["pl_id"], sg["pl_id"])
        for k in ("pl_rh", "pl_vh", "tp_rh", "tp_vh"):
            scale = np.max(np.abs(so[k])) if so[k].size else 1.0
            assert np.max(np.abs(so[k] - sg[k]), initial=0.0) <= 1e-10 * scale, (blk, k)
        assert np.array_equal(so["pl_status"], sg["pl_status"]) and np.array_equal(so["tp_status"], sg["tp_status"])
        deo = (_energy_of(so, sysic, integ, False) - e0o) / abs(e0o)
        deg = (_energy_of(sg, sysic, integ, True, g) - e0g) / abs(e0g)
        assert abs(deg - deo) <= 2e-12, (blk, deo, deg)
    if kind != "sparse":
        assert depth >= 2


# ---- discard checks (swg_system_discard) -----------------------------------
# Against the oracle after every step: the discard counts and every status
# exactly; the state (including the velocities the h2b conversion rewrites)
# to 1e-12 relative.  The systems hold unbound and very eccentric particles,
# whose drifts take the two branches that call libm transcendentals (see the
# module docstring), so their trajectories are not bit-identical.

def _discard_system(seed, npl=6, ntp=400, near=True):
    """Planets plus test particles spread so that each check fires: some far
    out, some close to the central body, some unbound, some on eccentric
    orbits with small pericentres and (near=True) some next to a planet.
    SyMBA resolves the latter as collisions inside its step, which the
    oracle leaves to the host, so its systems use near=False."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = ics.whm_solar(seed=seed, npl=npl, ntp=0)
    m = ntp
    a = np.exp(rng.uniform(np.log(0.15), np.log(60.0), m))
    e = np.where(rng.uniform(0, 1, m) < 0.3, rng.uniform(0.5, 0.95, m), rng.uniform(0, 0.1, m))
    inc = rng.uniform(0, 0.1, m)
    capom, omega, capm = (rng.uniform(0, 2 * np.pi, m) for _ in range(3))
    r, v = ics.el2xv(np.full(m, s.GMcb), a, e, inc, capom, omega, capm)
    v[: m // 10] *= 1.6                       # a tenth unbound
    k = m // 10 if near else 0
    j = rng.integers(0, npl, k)               # a tenth next to a planet
    if k:
        r[m - k:] = s.pl_rh[j] + rng.normal(0, 1e-3, (k, 3))
        v[m - k:] = s.pl_vh[j] + rng.normal(0, 0.05, (k, 3))
    s.tp_rh, s.tp_vh = r, v
    s.radius = s.radius * 100.0
    return s


@pytest.mark.parametrize("integ", [S.HELIO, S.SYMBA, S.WHM])
@pytest.mark.parametrize("qmin_coord", [0, 1])
def test_discard_matches_oracle(integ, qmin_coord):
    print(f"discard integ={integ} qmin_coord={qmin_coord}", flush=True)
    sysic = _discard_system(41 + integ, near=integ != S.SYMBA)
    kw = dict(rmin=0.2, rmax=40.0, rmaxu=30.0, qmin=0.3, qmin_alo=0.1, qmin_ahi=50.0, qmin_coord=qmin_coord)
    g = S.GpuSystem(integ, sysic, 0.01, lclose=True)
    o = O.OracleSystem(integ, sysic, 0.01, lclose=True)
    tot = [0, 0]
    for step in range(12):
        g.step()
        assert o.step() == 0
        got, ref = g.discard(**kw), o.discard(**kw)
        assert got == ref, (step, got, ref)
        tot[0] += got[0]
        tot[1] += got[1]
        sg, so = g.state(), o.state()
        sg["cb_vb"] = sg["cb_vb"][0]
        for k in ("tp_status", "pl_status"):
            assert np.array_equal(sg[k], so[k]), (step, k, np.argwhere(sg[k] != so[k])[:4].tolist())
        for k in ("pl_rh", "pl_vh", "tp_rh", "tp_vh", "cb_vb", "pl_vb", "tp_vb"):
            np.testing.assert_allclose(sg[k], so[k], rtol=1e-12, atol=1e-14, err_msg=f"step {step} {k}")
        if got[0]:
            break  # a system that lost massive bodies is rebuilt by the host
    assert tot[1] > 20  # the checks fired


def test_discard_symba_massive_bodies_and_ensemble():
    systems = [_discard_system(51, npl=8, ntp=100, near=False), _discard_system(52, npl=5, ntp=0, near=False),
               _discard_system(53, npl=3, ntp=60, near=False)]
    systems[0].pl_rh[2], systems[0].pl_vh[2] = (100.0, 0.0, 0.0), (0.0, 0.63, 0.0)  # beyond rmax
    systems[2].pl_rh[1], systems[2].pl_vh[1] = (35.0, 0.0, 0.0), (0.0, 2.0, 0.0)    # unbound, beyond rmaxu
    kw = dict(rmin=0.1, rmax=45.0, rmaxu=30.0)
    g = S.GpuSystem(S.SYMBA, systems, 0.01, lclose=True)
    g.step()
    got = g.discard(**kw)
    sg = g.state()
    po = to = 0
    tot = [0, 0]
    for s in systems:
        o = O.OracleSystem(S.SYMBA, s, 0.01, lclose=True)
        assert o.step() == 0
        c = o.discard(**kw)
        tot[0] += c[0]
        tot[1] += c[1]
        so = o.state()
        assert np.array_equal(sg["pl_status"][po:po + s.npl], so["pl_status"])
        assert np.array_equal(sg["tp_status"][to:to + s.ntp], so["tp_status"])
        np.testing.assert_allclose(sg["tp_vb"][to:to + s.ntp], so["tp_vb"], rtol=1e-12, atol=1e-14)
        po += s.npl
        to += s.ntp
    assert tuple(tot) == got and got[0] == 2


def test_failed_tile_allocation_is_an_error_not_a_launch(monkeypatch):
    """A failed allocation of the tile arrays (forced from the host with
    SWG_FAIL_TILE_ALLOC) comes back as SWG_ERR_HIP from set_frozen, and the
    system then refuses to step or time its kernels (no launch over a missing
    tile array) until a rebuild succeeds; after that it steps exactly as a
    system that never failed."""
    systems = [compact_disk(seed=30 + k, npl=n, ntp=t) for k, (n, t) in enumerate(((50, 20), (90, 0)))]
    ref = S.GpuSystem(S.SYMBA, systems, 0.005, lclose=False)
    g = S.GpuSystem(S.SYMBA, systems, 0.005, lclose=False)
    monkeypatch.setenv("SWG_FAIL_TILE_ALLOC", "1")
    with pytest.raises(S.SwiftestHipError) as e:
        g.set_frozen(np.array([0, 0], np.int32))
    assert e.value.code == -2 and "injected" in str(e.value)
    for call in (lambda: g.run(1), g.accel):
        with pytest.raises(S.SwiftestHipError) as e:
            call()
        assert e.value.code == -2 and "tile arrays are missing" in str(e.value)
    monkeypatch.delenv("SWG_FAIL_TILE_ALLOC")
    g.set_frozen(np.array([0, 0], np.int32))
    g.run(3)
    ref.run(3)
    a, b = ref.state(), g.state()
    for k in ("pl_rh", "pl_vb", "tp_rh", "tp_vb"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("seed,ntp", [(31, 40), (33, 0)])
def test_collision_stop_content_matches_the_oracle(seed, ntp):
    """Crowded disks that collide within 20 steps, stepped one at a time:
    the step in which the collision is found, the colliding pairs (flagged
    COLLIDED in the level-0 list), the recursion counts of that step and
    the state at its end equal the oracle's, which flags the pairs and
    finishes the step as the device does."""
    d = compact_disk(seed=seed, npl=300, ntp=ntp)
    o = O.OracleSystem(O.SYMBA, d, 0.005, collision_finish_step=True)
    g = S.GpuSystem(S.SYMBA, d, 0.005)
    for k in range(20):
        print(f"collision seed {seed} step {k}", flush=True)
        rc = o.run(1)
        try:
            g.run(1)
            grc = 0
        except S.SwiftestHipError as e:
            grc = e.code
        assert (rc == -1) == (grc == -3), (k, rc, grc)
        if rc:
            ncol = _compare_collision(o, g, f"seed {seed} step {k}")
            print(f"collision at step {k}: {ncol} pair(s) flagged", flush=True)
            # the run stays stopped: the next step reports it again
            with pytest.raises(S.SwiftestHipError) as e:
                g.run(1)
            assert e.value.code == -3
            return
        _compare_systems(o, g, f"seed {seed} step {k}")
    pytest.fail("no collision within 20 steps")
