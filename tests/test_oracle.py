"""CPU checks that pin the restatement (oracle/) itself.

The reference ships no golden vectors or known-answer outputs for this path,
so each oracle function is pinned by mathematics independent of its own
formulas: closed forms, exact rational arithmetic, analytic two-body motion,
conservation laws, and the reference test suite's conservation thresholds
(tests/test_swiftest.py:118-121).
"""
from fractions import Fraction

import numpy as np
import pytest

import pyoracle as O
from helpers import compact_disk
from swiftest_amd import ics


# ---------------------------------------------------------------- gravity --

def test_two_body_closed_form():
    r = np.array([[0.3, -0.2, 0.1], [1.3, 0.5, -0.4]])
    Gm = np.array([2e-4, 3e-3])
    acc = O.accel_plpl(r, Gm)
    d = r[1] - r[0]
    dist = np.sqrt(d @ d)
    np.testing.assert_allclose(acc[0], Gm[1] * d / dist ** 3, rtol=2e-16 * 8)
    np.testing.assert_allclose(acc[1], -Gm[0] * d / dist ** 3, rtol=2e-16 * 8)


def test_pair_sum_momentum_and_extended_precision():
    rng = np.random.Generator(np.random.PCG64(1))
    n = 64
    r = rng.normal(0, 1, (n, 3))
    Gm = rng.uniform(1e-6, 1e-3, n)
    acc = O.accel_plpl(r, Gm)
    # sum_i m_i a_i = 0 (Newton's third law), to rounding
    mom = (Gm[:, None] * acc).sum(0)
    scale = (Gm[:, None] * np.abs(acc)).sum(0)
    assert np.all(np.abs(mom) <= 1e-13 * scale)
    # against an extended-precision direct sum
    rl = r.astype(np.longdouble)
    ref = np.zeros((n, 3), np.longdouble)
    for i in range(n):
        d = rl - rl[i]
        d2 = (d * d).sum(1)
        d2[i] = 1
        f = Gm.astype(np.longdouble) / (d2 * np.sqrt(d2))
        f[i] = 0
        ref[i] = (f[:, None] * d).sum(0)
    np.testing.assert_allclose(acc, ref.astype(np.float64), rtol=1e-12, atol=1e-16 * np.abs(acc).max())


def test_radius_overlap_excluded():
    r = np.array([[0.0, 0.0, 0.0], [1e-3, 0.0, 0.0], [1.0, 0.0, 0.0]])
    Gm = np.array([1e-5, 1e-5, 1e-5])
    R = np.array([1e-3, 1e-3, 1e-4])   # bodies 0 and 1 overlap
    acc = O.accel_plpl(r, Gm, R)
    assert np.isclose(acc[0][0], Gm[2] / 1.0, rtol=1e-15)
    acc_n = O.accel_plpl(r, Gm)
    assert acc_n[0][0] > acc[0][0]


def test_pair_list_removal_equals_full_rows():
    rng = np.random.Generator(np.random.PCG64(2))
    n = 20
    r = rng.normal(0, 1, (n, 3))
    Gm = rng.uniform(1e-6, 1e-4, n)
    k = np.array([(i + 1, j + 1) for i in range(n) for j in range(i + 1, n)], np.int32)
    full = O.accel_plpl(r, Gm)
    pairs = O.accel_pairs(k, r, Gm)
    np.testing.assert_allclose(pairs, full, rtol=1e-12, atol=1e-15)


def test_tp_accel_extended_precision():
    rng = np.random.Generator(np.random.PCG64(3))
    rpl = rng.normal(0, 1, (30, 3))
    Gm = rng.uniform(1e-6, 1e-3, 30)
    rtp = rng.normal(0, 2, (50, 3))
    acc = O.accel_pltp(rtp, rpl, Gm)
    d = rtp[:, None, :].astype(np.longdouble) - rpl[None].astype(np.longdouble)
    d2 = (d * d).sum(-1)
    ref = -(Gm[None, :, None] * d / (d2 * np.sqrt(d2))[..., None]).sum(1)
    np.testing.assert_allclose(acc, ref.astype(np.float64), rtol=1e-13, atol=1e-18)


# ------------------------------------------------------------------ drift --

def _kepler_analytic(mu, r0, v0, dt):
    """Independent two-body propagation: elements, mean anomaly advance,
    Kepler's equation by bisection+Newton in extended precision, back to xv."""
    r0 = np.asarray(r0, np.longdouble)
    v0 = np.asarray(v0, np.longdouble)
    mu = np.longdouble(mu)
    rr = np.sqrt(r0 @ r0)
    a = 1 / (2 / rr - (v0 @ v0) / mu)
    n = np.sqrt(mu / a ** 3)
    ecosE = 1 - rr / a
    esinE = (r0 @ v0) / (n * a * a)
    E0 = np.arctan2(esinE, ecosE)
    e = np.sqrt(ecosE ** 2 + esinE ** 2)
    M = E0 - esinE + n * np.longdouble(dt)
    E = M
    for _ in range(100):
        E = E - (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
    dE = E - E0
    f = a / rr * (np.cos(dE) - 1) + 1
    g = np.longdouble(dt) + (np.sin(dE) - dE) / n
    r1 = f * r0 + g * v0
    r1m = np.sqrt(r1 @ r1)
    fd = -a * a * n * np.sin(dE) / (rr * r1m)
    gd = 1 + a / r1m * (np.cos(dE) - 1)
    return (r1).astype(np.float64), (fd * r0 + gd * v0).astype(np.float64)


@pytest.mark.parametrize("dt", [0.005, 0.3, 5.0])
def test_drift_matches_analytic_kepler(dt):
    rng = np.random.Generator(np.random.PCG64(4))
    a = rng.uniform(0.3, 20, 40)
    e = rng.uniform(0, 0.6, 40)
    inc = rng.uniform(0, 1, 40)
    r, v = ics.el2xv(ics.GMSUN, a, e, inc, *(rng.uniform(0, 2 * np.pi, 40) for _ in range(3)))
    x, vv, fl = O.drift(ics.GMSUN, r, v, dt)
    assert not fl.any()
    for i in range(len(r)):
        ra, va = _kepler_analytic(ics.GMSUN, r[i], v[i], dt)
        np.testing.assert_allclose(x[i], ra, rtol=0, atol=1e-11 * np.linalg.norm(ra))
        np.testing.assert_allclose(vv[i], va, rtol=0, atol=1e-11 * np.linalg.norm(va))


def test_drift_conserves_energy_and_momentum():
    rng = np.random.Generator(np.random.PCG64(5))
    r, v = ics.el2xv(ics.GMSUN, rng.uniform(0.5, 5, 200), rng.uniform(0, 0.9, 200), rng.uniform(0, 1, 200),
                     *(rng.uniform(0, 2 * np.pi, 200) for _ in range(3)))
    v = v * rng.uniform(0.5, 1.8, (200, 1))  # includes hyperbolic orbits (Laguerre / cubic guess)
    x, vv, fl = O.drift(ics.GMSUN, r, v, 0.7)
    ok = fl == 0
    E0 = 0.5 * (v * v).sum(1) - ics.GMSUN / np.linalg.norm(r, axis=1)
    E1 = 0.5 * (vv * vv).sum(1) - ics.GMSUN / np.linalg.norm(x, axis=1)
    np.testing.assert_allclose(E1[ok], E0[ok], rtol=1e-11, atol=1e-12)
    L0, L1 = np.cross(r, v), np.cross(x, vv)
    np.testing.assert_allclose(L1[ok], L0[ok], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("x", [1e-3, 0.05, 0.7, 3.0, 40.0, -0.3, -5.0, -30.0])
def test_stumpff_closed_forms(x):
    c0, c1, c2, c3 = O.stumpff(x)
    s = np.sqrt(abs(x))
    if x > 0:
        e0, e1, e2, e3 = np.cos(s), np.sin(s) / s, (1 - np.cos(s)) / x, (s - np.sin(s)) / (x * s)
    else:
        e0, e1, e2, e3 = np.cosh(s), np.sinh(s) / s, (np.cosh(s) - 1) / -x, (np.sinh(s) - s) / (-x * s)
    for a, b in ((c0, e0), (c1, e1), (c2, e2), (c3, e3)):
        assert abs(a - b) <= 1e-13 * max(1.0, abs(b))


# -------------------------------------------------------------- encounters --

def _exact_encounter(xr, vr, R, dt):
    """Closed-form minimum distance of linear relative motion over [0, dt] in
    exact rational arithmetic, with the approach condition v.r < 0."""
    x = [Fraction(float(q)) for q in xr]
    v = [Fraction(float(q)) for q in vr]
    R2 = Fraction(float(R)) ** 2
    r2 = sum(q * q for q in x)
    vdotr = sum(a * b for a, b in zip(x, v))
    v2 = sum(q * q for q in v)
    if r2 <= R2:
        return True, 0
    if vdotr >= 0 or v2 == 0:
        return False, abs(float(r2 - R2))
    tmin = -vdotr / v2
    t = min(tmin, Fraction(float(dt)))
    rmin2 = r2 + 2 * vdotr * t + v2 * t * t
    return rmin2 <= R2, abs(float((rmin2 - R2) / R2))


def test_encounter_branches():
    # inside the critical radius: encounter whatever the velocity
    assert O.encounter_one((0.1, 0, 0), (1, 0, 0), 0.5, 0.01) == (True, True)
    # receding: no encounter
    assert O.encounter_one((2, 0, 0), (1, 0, 0), 1.0, 10.0) == (False, False)
    # closest approach inside the step (tmin < dt)
    assert O.encounter_one((2, 0.5, 0), (-1, 0, 0), 0.6, 5.0) == (True, True)
    # approaching but the step ends first (tmin >= dt), endpoint distance 1.0 == R: inclusive
    assert O.encounter_one((2, 0, 0), (-1, 0, 0), 1.0, 1.0) == (True, True)
    assert O.encounter_one((2, 0, 0), (-1, 0, 0), 0.999, 1.0) == (False, True)
    # no relative motion
    assert O.encounter_one((2, 0, 0), (0, 0, 0), 1.0, 1.0) == (False, False)


def test_encounter_list_exact_rational_brute_force():
    rng = np.random.Generator(np.random.PCG64(6))
    n = 60
    r = rng.uniform(-1, 1, (n, 3))
    v = rng.normal(0, 1, (n, 3))
    renc = rng.uniform(0.02, 0.15, n)
    dt = 0.1
    i1, i2, lv = O.encounters_plpl(r, v, renc, dt)
    got = set(zip(i1.tolist(), i2.tolist()))
    assert list(zip(i1, i2)) == sorted(got)   # ordered index1 then index2
    for i in range(n):
        for j in range(i + 1, n):
            enc, margin = _exact_encounter(r[j] - r[i], v[j] - v[i], renc[i] + renc[j], dt)
            if margin < 1e-12:
                continue
            assert ((i + 1, j + 1) in got) == enc, (i, j)
    assert lv.all()


@pytest.mark.parametrize("nplm", [45, 20, 0])
def test_encounter_list_plplm_exact_rational_brute_force(nplm):
    """encounter_check_all_plplm (encounter_check.f90:42-109): every pair
    among the first nplm bodies and every pair of one of them with the rest;
    none among the rest; listed index1 then index2 ascending."""
    rng = np.random.Generator(np.random.PCG64(16))
    n = 60
    r = rng.uniform(-1, 1, (n, 3))
    v = rng.normal(0, 1, (n, 3))
    renc = rng.uniform(0.02, 0.15, n)
    dt = 0.1
    i1, i2, lv = O.encounters_plplm(r, v, renc, nplm, dt)
    got = set(zip(i1.tolist(), i2.tolist()))
    assert list(zip(i1, i2)) == sorted(got)
    for i in range(n):
        for j in range(i + 1, n):
            enc, margin = _exact_encounter(r[j] - r[i], v[j] - v[i], renc[i] + renc[j], dt)
            if margin < 1e-12:
                continue
            assert ((i + 1, j + 1) in got) == (enc and i < nplm), (i, j)
    assert lv.all()


@pytest.mark.parametrize("q", [0.5, 0.75])
def test_semi_interacting_bodies_in_the_symba_step(q):
    """GMTINY: the bodies below it feel the others and move the others, but
    not each other.  Over 100 SyMBA steps of a disk whose lighter half (or
    three quarters: the triangular branch of swiftest_kick.f90:190-216) is
    below GMTINY, their pairs never enter the encounter list, and the
    accelerations differ from the fully interacting ones by exactly the tiny
    bodies' mutual attraction."""
    d = compact_disk(seed=23, npl=60, mlo=1e-9, mhi=3e-6, alo=0.9, ahi=1.1)
    gmt = float(np.quantile(d.Gmass, q))
    o = O.OracleSystem(O.SYMBA, d, 0.005, gmtiny=gmt)
    st = o.state()
    Gm = d.Gmass[st["pl_id"] - 1]
    nplm = int((Gm >= gmt).sum())
    assert 0 < nplm < len(Gm)
    for _ in range(20):
        assert o.run(5) == 0
        i1, i2, _ = o.enc_list(0)
        assert np.all(i1 <= nplm)          # index1 always fully interacting
    # the tiny bodies' mutual forces are absent: the tiny-tiny pair sum of a
    # step equals the difference of the two kicks' sums
    r = st["pl_rh"]
    a_full = O.accel_plpl(r, Gm)
    a_tiny = O.accel_plpl(r, Gm, nplm=nplm)
    k = np.arange(len(Gm)) >= nplm
    mutual = np.zeros_like(a_full)
    for i in np.flatnonzero(k):
        for j in np.flatnonzero(k):
            if i != j:
                dr = r[j] - r[i]
                mutual[i] += Gm[j] * dr / np.linalg.norm(dr) ** 3
    np.testing.assert_allclose(a_full - a_tiny, mutual, rtol=1e-9, atol=1e-14 * np.abs(a_full).max())


# ---------------------------------------------------------- shell kicks --

def test_pow_m1p5_correctly_rounded():
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    rng = np.random.Generator(np.random.PCG64(7))
    xs = np.exp(rng.uniform(-20, 10, 3000))
    for x in xs:
        got = O.pow_m1p5(x)
        exact = Decimal(float(x)) ** Decimal(-1.5)
        lo, hi = np.nextafter(got, 0), np.nextafter(got, np.inf)
        # the exact value is closer to got than to either neighbour
        assert abs(Decimal(float(got)) - exact) <= abs(Decimal(float(lo)) - exact)
        assert abs(Decimal(float(got)) - exact) <= abs(Decimal(float(hi)) - exact)
        assert abs(got - x ** -1.5) <= np.spacing(got)   # within 1 ulp of libm pow


def test_powi():
    # the shell factor itself is pinned in test_oracle_symba_pins.py
    for n in range(0, 40):
        assert O.powi(0.48075, n) == pytest.approx(0.48075 ** n, rel=1e-15)
    assert O.powi(0.48075, 0) == 1.0


# ------------------------------------------------------------ whole runs --

def _energy(GMcb, rh, vh, Gm):
    Mt = GMcb + Gm.sum()
    vcb = -(Gm[:, None] * vh).sum(0) / Mt
    vb = vh + vcb
    ke = 0.5 * (Gm * (vb * vb).sum(1)).sum() + 0.5 * GMcb * (vcb @ vcb)
    pe = -(GMcb * Gm / np.linalg.norm(rh, axis=1)).sum()
    d = np.linalg.norm(rh[:, None] - rh[None], axis=-1)
    iu = np.triu_indices(len(Gm), 1)
    pe -= (Gm[:, None] * Gm[None] / np.where(d > 0, d, np.inf))[iu].sum()
    # barycentric angular momentum; the central body's and the frame-offset
    # terms cancel because the total barycentric momentum is zero
    Lv = (Gm[:, None] * np.cross(rh, vb)).sum(0)
    return ke + pe, np.linalg.norm(Lv)


@pytest.mark.parametrize("integ", [O.WHM, O.HELIO, O.SYMBA])
def test_conservation_thresholds(integ):
    """The reference's own invariant test (tests/test_swiftest.py:112-170):
    |dE/E0| slope < 1e-8 /yr and |dL/L0| slope < 1e-10 /yr, here on an
    offline synthetic 8-planet system over 200 yr."""
    # the eight planets' masses (M_sun), semi-major axes (AU), eccentricities
    # and inclinations (deg), rounded; orientation angles drawn with a fixed seed
    m = np.array([1.66e-7, 2.45e-6, 3.0e-6, 3.2e-7, 9.55e-4, 2.86e-4, 4.37e-5, 5.15e-5])
    a = np.array([0.387, 0.723, 1.0, 1.524, 5.203, 9.537, 19.19, 30.07])
    e = np.array([0.2056, 0.0068, 0.0167, 0.0934, 0.0489, 0.0565, 0.0457, 0.0113])
    inc = np.radians([7.0, 3.39, 0.0, 1.85, 1.30, 2.49, 0.77, 1.77])
    rng = np.random.Generator(np.random.PCG64(7))
    Gm = ics.GU * m
    r, v = ics.el2xv(ics.GMSUN + Gm, a, e, inc, *(rng.uniform(0, 2 * np.pi, 8) for _ in range(3)))
    s = ics.System(ics.GMSUN, ics.RSUN_AU, r, v, Gm, ics._radius_from_mass(m), np.zeros((0, 3)), np.zeros((0, 3)))
    o = O.OracleSystem(integ, s, 0.01, istep_out=100)
    st = o.state()
    Gm = s.Gmass[st["pl_id"] - 1]
    E0, L0 = _energy(s.GMcb, st["pl_rh"], st["pl_vh"], Gm)
    ts, dE, dL = [], [], []
    for k in range(20):
        assert o.run(1000) == 0
        st = o.state()
        E, L = _energy(s.GMcb, st["pl_rh"], st["pl_vh"], Gm)
        ts.append((k + 1) * 10.0)
        dE.append(abs((E - E0) / E0))
        dL.append(abs((L - L0) / L0))
    assert abs(np.polyfit(ts, dE, 1)[0]) < 1e-8
    assert abs(np.polyfit(ts, dL, 1)[0]) < 1e-10


def test_symba_disk_energy_bounded():
    d = compact_disk(seed=9, npl=80, mhi=3e-6, alo=0.9, ahi=1.1)
    o = O.OracleSystem(O.SYMBA, d, 0.005, istep_out=20)
    st = o.state()
    Gm = d.Gmass[st["pl_id"] - 1]
    E0, L0 = _energy(d.GMcb, st["pl_rh"], st["pl_vh"], Gm)
    assert o.run(200) == 0
    assert o.stats()["max_depth"] >= 0
    st = o.state()
    E, L = _energy(d.GMcb, st["pl_rh"], st["pl_vh"], Gm)
    # the oracle reaches 2.7e-11; the bound is about 100x that (a dropped
    # shell taper gives 6.5e-8, dropped -1 shell kicks 1.1e-7)
    assert abs((E - E0) / E0) < 3e-9
    assert abs((L - L0) / L0) < 1e-10


def test_short_circuit_skips_tp_check_when_plpl_encounters():
    d = compact_disk(seed=10, npl=60, ntp=40)
    on = O.OracleSystem(O.SYMBA, d, 0.005, short_circuit=True)
    off = O.OracleSystem(O.SYMBA, d, 0.005, short_circuit=False)
    on.step()
    off.step()
    assert on.stats()["nplpl_enc"] > 0
    assert on.stats()["npltp_enc"] == 0
    assert off.stats()["npltp_enc"] > 0


# ----------------------------------------------------------------- energy --
# swo_energy restates swiftest_util_get_energy_and_momentum_system
# (swiftest_util.f90:1172-1291) after swiftest_util_coord_rh2rb_pl (:488-518).
# The reference takes the central-body potential of body i at its barycentric
# distance |rb_i| (swiftest_util.f90:1368), so the pins below use
# configurations where the facts they check hold exactly.

def test_norm2_is_the_euclidean_norm_without_overflow():
    rng = np.random.Generator(np.random.PCG64(11))
    for _ in range(2000):
        x, y, z = rng.normal(0, 1, 3) * 10.0 ** rng.integers(-5, 5, 3)
        s2 = Fraction(x) ** 2 + Fraction(y) ** 2 + Fraction(z) ** 2  # exact
        r = O.norm2(x, y, z)
        ulp = np.spacing(r)
        assert Fraction(r - 4 * ulp) ** 2 <= s2 <= Fraction(r + 4 * ulp) ** 2  # scaled sum: a few ulps
    assert O.norm2(3e200, 4e200, 0.0) == pytest.approx(5e200, rel=1e-15)  # scaled: no overflow
    assert O.norm2(0.0, 0.0, 0.0) == 0.0
    assert O.norm2(-2.0, 0.0, 0.0) == 2.0


def _square(a=1.3, u=5.1, Gm=2e-3, GMcb=ics.GMSUN, GU=ics.GU):
    """Four equal bodies on a square around the central body, each moving
    tangentially at speed u; the central body is at rest.  The positions sum
    to zero exactly, so the barycentre is the central body."""
    rh = np.array([[a, 0, 0], [0, a, 0], [-a, 0, 0], [0, -a, 0]], float)
    vb = np.array([[0, u, 0], [-u, 0, 0], [0, -u, 0], [u, 0, 0]], float)
    return rh, vb, np.full(4, Gm), np.full(4, Gm / GU), GMcb / GU


def test_energy_closed_form_square():
    a, u, Gm = 1.3, 5.1, 2e-3
    rh, vb, G, m, mcb = _square(a, u, Gm)
    out, rb = O.energy(ics.GMSUN, mcb, ics.RSUN_AU, np.zeros(3), rh, vb, G, m)
    assert np.array_equal(rb, rh)
    pe = -4 * ics.GMSUN * m[0] / a - Gm * m[0] * (4 / (a * np.sqrt(2)) + 2 / (2 * a))
    assert out[O_E["pe"]] == pytest.approx(pe, rel=4e-16 * 8)
    assert out[O_E["ke_orbit"]] == pytest.approx(0.5 * 4 * m[0] * u * u, rel=4e-16 * 4)
    np.testing.assert_allclose(out[6:9], [0, 0, 4 * m[0] * a * u], rtol=4e-16 * 4, atol=0)
    assert out[O_E["te"]] == pytest.approx(out[O_E["ke_orbit"]] + out[O_E["pe"]], rel=1e-16)
    assert out[O_E["GMtot"]] == pytest.approx(ics.GMSUN + 4 * Gm, rel=1e-16)
    assert out[O_E["be_cb"]] == pytest.approx(-3 * ics.GMSUN * mcb / (5 * ics.RSUN_AU), rel=1e-16)
    np.testing.assert_array_equal(out[12:15], out[6:9])
    assert out[O_E["ke_spin"]] == 0 and out[O_E["be"]] == 0


def test_energy_barycentre_and_frame_invariance():
    rng = np.random.Generator(np.random.PCG64(12))
    n = 40
    rh = rng.normal(0, 2, (n, 3))
    rh[0] = [1e-3, -2e-3, 5e-4]  # close to the central body: rb[0] - rh[0] is the offset to ~1e-19
    vb = rng.normal(0, 3, (n, 3))
    Gm = ics.GU * np.exp(rng.uniform(np.log(1e-7), np.log(1e-3), n))
    m = Gm / ics.GU
    cb_vb = rng.normal(0, 1e-3, 3)
    out, rb = O.energy(ics.GMSUN, 1.0, ics.RSUN_AU, cb_vb, rh, vb, Gm, m)
    off = rb - rh                                      # one offset for every body
    assert np.allclose(off, off[0], rtol=0, atol=1e-15)
    bary = (Gm[:, None] * rb).sum(0) + ics.GMSUN * off[0]
    assert np.all(np.abs(bary) <= 1e-14 * ((Gm[:, None] * np.abs(rb)).sum(0) + ics.GMSUN * np.abs(off[0])))
    # every pair once, i < j: the pair part is -sum_{i<j} Gm_i m_j / |rb_i - rb_j|
    pairs = sum(-Gm[i] * m[j] / np.linalg.norm(rb[i] - rb[j]) for i in range(n) for j in range(i + 1, n))
    cb = sum(-ics.GMSUN * m[i] / np.linalg.norm(rb[i]) for i in range(n))
    assert out[O_E["pe"]] == pytest.approx(pairs + cb, rel=1e-13)
    ke = 0.5 * (1.0 * cb_vb @ cb_vb + sum(m[i] * vb[i] @ vb[i] for i in range(n)))
    assert out[O_E["ke_orbit"]] == pytest.approx(ke, rel=1e-13)
    L = np.cross(off[0], cb_vb) * 1.0 + sum(m[i] * np.cross(rb[i], vb[i]) for i in range(n))
    assert np.linalg.norm(out[6:9] - L) <= 1e-13 * np.linalg.norm(L)


def test_energy_inactive_bodies_are_left_out_exactly():
    rng = np.random.Generator(np.random.PCG64(13))
    n = 30
    rh, vb = rng.normal(0, 2, (n, 3)), rng.normal(0, 3, (n, 3))
    Gm = ics.GU * rng.uniform(1e-7, 1e-4, n)
    R = rng.uniform(1e-5, 1e-4, n)
    st = np.zeros(n, np.int32)
    st[[3, 17, 29]] = 1  # INACTIVE
    keep = st == 0
    cbv = np.array([1e-4, -2e-4, 3e-5])
    a, _ = O.energy(ics.GMSUN, 1.0, ics.RSUN_AU, cbv, rh, vb, Gm, Gm / ics.GU, R, st, lclose=True)
    b, _ = O.energy(ics.GMSUN, 1.0, ics.RSUN_AU, cbv, rh[keep], vb[keep], Gm[keep], Gm[keep] / ics.GU, R[keep],
                    lclose=True)
    assert np.array_equal(a, b)
    be = sum(-3 * Gm[i] * (Gm[i] / ics.GU) / (5 * R[i]) for i in range(n) if keep[i])
    assert a[O_E["be"]] == pytest.approx(be, rel=1e-14)
    assert a[O_E["te"]] == pytest.approx(a[0] + a[2] + a[3], rel=1e-16)


O_E = {"ke_orbit": 0, "ke_spin": 1, "pe": 2, "be": 3, "be_cb": 4, "te": 5, "GMtot": 15}


# --------------------------------------------------------------- discards --
# swo_discard restates swiftest_discard_system and its callees
# (swiftest_discard.f90:13-337, symba_discard.f90:14-387).  Pinned by planted
# bodies on either side of each threshold, the exact minimum distance of
# linear relative motion for the close-approach test, a Kepler orbit's known
# pericentre for the passage test, the reference's pericentre-flag sequence,
# and zero barycentric momentum after the h2b conversion it performs.

def _dstate(tp_rh, tp_vh, integ=O.HELIO, pl_rh=((3.0, 0, 0),), pl_vh=((0, 3.6, 0),), Gm=(1e-3,), R=(1e-3,),
            rhill=(0.2,)):
    pl_rh, pl_vh = np.array(pl_rh, float), np.array(pl_vh, float)
    return O.DiscardState(integ, ics.GMSUN, 0.005, np.zeros(3), pl_rh, pl_vh, pl_vh.copy(), np.array(Gm, float),
                          np.array(R, float), np.array(rhill, float), np.array(tp_rh, float), np.array(tp_vh, float),
                          np.array(tp_vh, float))


def test_discard_central_body_thresholds():
    tp_rh = [[10.0 - 1e-12, 0, 0], [10.0 + 1e-12, 0, 0], [0.05, 0, 0], [0.003, 0, 0], [60.0, 0, 0], [60.0, 0, 0]]
    esc = np.sqrt(2 * ics.GMSUN / 60.0)
    tp_vh = [[0, 2, 0], [0, 2, 0], [0, 20, 0], [0, 20, 0], [0, 1.1 * esc, 0], [0, 0.9 * esc, 0]]
    st = _dstate(tp_rh, tp_vh)
    st.Rcb = 0.004  # rmin 0.01 floored at the central body's radius only when larger
    n = st.discard(0.01, rmin=0.01, rmax=100.0, rmaxu=50.0)
    assert list(st.tp_status) == [0, 0, 0, -2, -3, 0] and n == (0, 2)
    st = _dstate(tp_rh, tp_vh)
    n = st.discard(0.01, rmax=10.0)
    assert list(st.tp_status) == [0, -1, 0, 0, -1, -1] and n == (0, 3)
    st = _dstate(tp_rh, tp_vh)
    st.Rcb = 0.06  # cb radius above rmin: the floor is the radius
    st.discard(0.01, rmin=0.01)
    assert list(st.tp_status) == [0, 0, -2, -2, 0, 0]


def test_discard_h2b_zero_barycentric_momentum():
    rng = np.random.Generator(np.random.PCG64(21))
    n = 12
    pl_rh, pl_vh = rng.normal(0, 3, (n, 3)), rng.normal(0, 4, (n, 3))
    Gm = ics.GU * rng.uniform(1e-6, 1e-3, n)
    st = _dstate(rng.normal(0, 5, (30, 3)), rng.normal(0, 3, (30, 3)), pl_rh=pl_rh, pl_vh=pl_vh, Gm=Gm,
                 R=np.full(n, 1e-4), rhill=np.full(n, 0.1))
    st.discard(0.01, rmax=1e9)
    mom = (Gm[:, None] * st.pl_vb).sum(0) + ics.GMSUN * st.cb_vb
    assert np.all(np.abs(mom) <= 1e-14 * (Gm[:, None] * np.abs(st.pl_vb)).sum(0).max())
    np.testing.assert_array_equal(st.tp_vb, st.tp_vh + st.cb_vb)
    np.testing.assert_array_equal(st.pl_vb, st.pl_vh + st.cb_vb)


@pytest.mark.parametrize("case", range(6))
def test_discard_close_approach_exact(case):
    # particle moving on a straight line past a body of radius R at the origin
    R, dt = 0.01, 0.5
    x0, v = [(0.005, 0, 0), (0.3, 0.02, 0), (0.3, 0.02, 0), (0.3, 0.005, 0), (0.3, 0.009, 0), (-0.3, 0.005, 0)][case], \
        [(0, 1, 0), (-1, 0, 0), (-0.5, 0, 0), (-0.1, 0, 0), (-1, 0, 0), (-1, 0, 0)][case]
    st = _dstate([x0], [v], pl_rh=[(0, 0, 0)], pl_vh=[(0, 0, 0)], R=[R])
    st.discard(dt, lclose=True)
    # exact minimum over [0, dt] of |x0 + v t|^2 (the reference evaluates the
    # minimum of the line when it falls inside dt, else the end point)
    X, V = [Fraction(q) for q in x0], [Fraction(q) for q in v]
    r2 = sum(q * q for q in X)
    vdotr = sum(a * b for a, b in zip(X, V))
    v2 = sum(q * q for q in V)
    if r2 <= Fraction(R) ** 2:
        hit = True
    elif vdotr > 0:
        hit = False
    else:
        tmin = -vdotr / v2
        T = tmin if tmin < Fraction(dt) else Fraction(dt)
        hit = min(sum((a + b * T) ** 2 for a, b in zip(X, V)), r2) <= Fraction(R) ** 2
    assert (st.tp_status[0] == -5) == hit, (case, st.tp_status[0])
    assert [True, False, False, False, True, False][case] == hit  # both outcomes covered


def test_discard_pericentre_sequence_and_value():
    # a particle moving outward right after pericentre of a known ellipse
    mu, a, e = ics.GMSUN, 1.5, 0.8
    q = a * (1 - e)
    r, v = ics.el2xv(np.array([mu]), np.array([a]), np.array([e]), np.zeros(1), np.zeros(1), np.zeros(1),
                     np.array([0.01]))
    st = _dstate(r, v, pl_rh=[(30.0, 0, 0)], pl_vh=[(0, 1.1, 0)])
    kw = dict(qmin=0.5, qmin_alo=1.0, qmin_ahi=2.0)
    st.discard(0.01, **kw)          # isperi 1 -> -1 (moving outward)
    assert st.tp_isperi[0] == -1 and st.tp_status[0] == 0
    st.discard(0.01, **kw)          # -1 and vdotr >= 0 -> 0: a and q formed, q <= qmin -> discarded
    assert st.tp_isperi[0] == 0 and st.tp_status[0] == -4
    assert st.tp_peri[0] == pytest.approx(q, rel=1e-12) and st.tp_atp[0] == pytest.approx(a, rel=1e-12)
    st2 = _dstate(r, v, pl_rh=[(r[0, 0], r[0, 1] + 0.1, 0)], pl_vh=[(0, 1.1, 0)], rhill=(0.2,))
    st2.discard(0.01, **kw)
    st2.discard(0.01, **kw)         # inside a massive body's Hill sphere: kept
    assert st2.tp_isperi[0] == 0 and st2.tp_status[0] == 0
    st3 = _dstate(r, -v, pl_rh=[(30.0, 0, 0)], pl_vh=[(0, 1.1, 0)])
    st3.discard(0.01, **kw)         # moving inward: isperi 1 -> 1
    assert st3.tp_isperi[0] == 1


def test_discard_symba_massive_bodies():
    pl_rh = [(3.0, 0, 0), (120.0, 0, 0), (0.01, 0, 0), (4.0, 0, 0)]
    pl_vh = [(0, 3.6, 0), (0, 0.5, 0), (0, 60, 0), (0, 3.1, 0)]
    st = _dstate([(5.0, 0, 0)], [(0, 2.8, 0)], integ=O.SYMBA, pl_rh=pl_rh, pl_vh=pl_vh, Gm=(1e-4,) * 4, R=(1e-4,) * 4,
                 rhill=(0.05,) * 4)
    n = st.discard(0.01, rmin=0.05, rmax=100.0)
    assert list(st.pl_status) == [0, -1, -2, 0] and n == (2, 0)
    # the test particles then see only the bodies left: h2b over bodies 0 and 3
    g = ics.GMSUN + 2e-4
    cb = -(1e-4 * np.array(pl_vh[0]) + 1e-4 * np.array(pl_vh[3])) / g
    np.testing.assert_allclose(st.cb_vb, cb, rtol=1e-15)


def test_collision_modes_agree_on_the_colliding_step():
    """A crowded disk that collides: stopping at the collision check and
    finishing the step report the collision in the same step, from the same
    state; finishing flags at least the pairs the stop flagged."""
    from helpers import compact_disk
    d = compact_disk(seed=33, npl=300, ntp=0)
    stop = O.OracleSystem(O.SYMBA, d, 0.005)
    fin = O.OracleSystem(O.SYMBA, d, 0.005, collision_finish_step=True)
    for k in range(20):
        a, b = stop.state(), fin.state()
        assert np.array_equal(a["pl_rh"], b["pl_rh"]) and np.array_equal(a["pl_vb"], b["pl_vb"]), k
        ra, rb = stop.run(1), fin.run(1)
        assert ra == rb, (k, ra, rb)
        if ra:
            assert ra == -1
            sa, sb = stop.enc_status(0), fin.enc_status(0)
            assert np.array_equal(stop.enc_list(0)[0], fin.enc_list(0)[0])
            assert (sa == -14).sum() >= 1
            assert np.all(sb[sa == -14] == -14)
            return
    pytest.fail("no collision within 20 steps")
