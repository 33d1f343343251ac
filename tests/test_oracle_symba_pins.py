"""Pins of the oracle's SyMBA recursion that do not go through the oracle's
own formulas:

- the shell factor of symba_kick_list_plpl / _pltp (symba_kick.f90:187-208)
  against an independent 50-digit evaluation, inside the taper and on both
  sides of both shell radii;
- a two-body hyperbolic flyby through the shells (a nearly massless central
  body, so that the relative motion is the Kepler problem) against the
  analytic hyperbola: relative energy, angular momentum and the
  Laplace-Runge-Lenz vector before and after the flyby, and the final
  relative position;
- planted straight-line encounters (2 bodies, a 3-body chain through one
  shared body, a 4-body chain, a massive body and a test particle) whose
  recursion levels and substep counts are derived by hand from the shell
  radii and the ternary substep grid (symba_step.f90:118-262,
  symba_encounter_check.f90:90-235): the expected values are literals;
- the planted-mistake builds of the oracle (a dropped taper polynomial, the
  -1 shell kicks left out) each fail at least one of these pins.
"""
from __future__ import annotations

from decimal import Decimal, getcontext

import numpy as np
import pytest

import pyoracle as O
from helpers import compact_disk
from swiftest_amd import ics

RSHELL = 0.48075
RHSCALE = 6.5
EPS = 2.0 ** -52


# ------------------------------------------------------------ shell factor --

def _shell_exact(ri, r2):
    """fac of symba_kick.f90:197-206 in 50-digit decimal arithmetic from the
    double inputs: r^-3 (1 - 3 rr^2 + 2 rr^3), rr = (sqrt(ri) - r) /
    (sqrt(ri) (1 - RSHELL)) inside the shell, r^-3 outside."""
    getcontext().prec = 50
    R2, RI = Decimal(float(r2)), Decimal(float(ri))
    r = R2.sqrt()
    inv3 = 1 / (R2 * r)
    if R2 >= RI:
        return inv3
    ris = RI.sqrt()
    rr = (ris - r) / (ris * (1 - Decimal(RSHELL)))
    return inv3 * (1 - 3 * rr * rr + 2 * rr * rr * rr)


def shell_factor_errors(L=None, n=400, seed=3):
    """Largest error of the oracle's shell factor, in units of eps * r^-3,
    over shell radii from 1e-6 to 10 AU and separations inside the taper,
    outside it, and one ulp either side of both shell radii; plus the list
    of boundary cases the oracle got wrong (kept / dropped)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    worst, wrong = 0.0, []
    for _ in range(n):
        ri = float(np.exp(rng.uniform(np.log(1e-12), np.log(100.0))))
        rim1 = ri * (RSHELL * RSHELL)
        inside = [float(rim1 + (ri - rim1) * u) for u in rng.uniform(0, 1, 4)]
        edges = [rim1, np.nextafter(rim1, np.inf), np.nextafter(ri, 0.0), ri, np.nextafter(ri, np.inf),
                 ri * 1.7]
        for r2 in inside + [float(x) for x in edges]:
            kept, fac = O.shell_fac(ri, r2, L)
            if not kept:
                wrong.append(("dropped", ri, r2))
                continue
            exact = _shell_exact(ri, r2)
            scale = Decimal(float(r2)) ** Decimal(-1.5)
            err = float(abs(Decimal(fac) - exact) / scale) / EPS
            worst = max(worst, err)
        below = float(np.nextafter(rim1, 0.0))
        kept, fac = O.shell_fac(ri, below, L)
        if kept:
            wrong.append(("kept", ri, below))
    return worst, wrong


def test_shell_factor_against_exact_evaluation():
    worst, wrong = shell_factor_errors()
    assert not wrong
    # the polynomial has |d/drr| <= 1.5 and rr is formed with a few
    # roundings, so the factor is good to a few eps of r^-3
    assert worst < 16.0


def test_shell_factor_continuity_at_both_radii():
    for ri in (1e-8, 3.7e-3, 0.25, 40.0):
        rim1 = ri * (RSHELL * RSHELL)
        # at the outer radius the taper is 1 and meets 1 / (r2 sqrt r2)
        _, f_in = O.shell_fac(ri, float(np.nextafter(ri, 0.0)))
        _, f_out = O.shell_fac(ri, ri)
        assert abs(f_in - f_out) <= 8 * EPS * f_out
        # at the inner radius the taper is 0
        _, f0 = O.shell_fac(ri, rim1)
        assert abs(f0) <= 8 * EPS * float(Decimal(rim1) ** Decimal(-1.5))


# ------------------------------------------------------- two-body flyby --

def _hyperbola(mu, r0, v0, t):
    """Analytic propagation of a hyperbolic Kepler orbit over t: hyperbolic
    anomaly from e sinh H - H = M by Newton in extended precision, then the
    f and g functions."""
    r0 = np.asarray(r0, np.longdouble)
    v0 = np.asarray(v0, np.longdouble)
    mu = np.longdouble(mu)
    r = np.sqrt(r0 @ r0)
    a = 1 / (2 / r - (v0 @ v0) / mu)
    assert a < 0
    n = np.sqrt(mu / (-a) ** 3)
    ech = 1 - r / a
    esh = (r0 @ v0) / np.sqrt(-mu * a)
    e = np.sqrt(ech * ech - esh * esh)
    H0 = np.arcsinh(esh / e)
    M = e * np.sinh(H0) - H0 + n * np.longdouble(t)
    H = np.arcsinh(M / e)
    for _ in range(200):
        H = H - (e * np.sinh(H) - H - M) / (e * np.cosh(H) - 1)
    dH = H - H0
    f = 1 - (a / r) * (1 - np.cosh(dH))
    g = np.longdouble(t) - (np.sinh(dH) - dH) / n
    r1 = f * r0 + g * v0
    r1m = np.sqrt(r1 @ r1)
    fd = -np.sqrt(-mu * a) * np.sinh(dH) / (r * r1m)
    gd = 1 - (a / r1m) * (1 - np.cosh(dH))
    return r1.astype(np.float64), (fd * r0 + gd * v0).astype(np.float64)


def _invariants(mu, r, v):
    r = np.asarray(r, np.longdouble)
    v = np.asarray(v, np.longdouble)
    rm = np.sqrt(r @ r)
    E = (v @ v) / 2 - mu / rm
    Lv = np.cross(r, v)
    A = np.cross(v, Lv) / mu - r / rm
    return float(E), Lv.astype(np.float64), A.astype(np.float64)


FLYBY_GM = 0.04        # each body, ~1e-3 M_sun
FLYBY_DT = 0.05
FLYBY_STEPS = 400


def flyby(L=None):
    """Two bodies start 10 AU apart and pass at impact parameter 0.05 AU
    with 1 AU/yr (pericentre 0.014 AU), eight shells inside their encounter
    radius (Hill radii 0.5 AU given, renc sum 6.5 AU); the central body's GM
    is 1e-14, so their relative motion is a Kepler hyperbola with
    mu = 0.08.  Returns the relative changes of the pair's energy, angular
    momentum and Laplace-Runge-Lenz vector over the flyby, the final
    relative position error against the hyperbola, and the deepest level."""
    rA, vA = np.array([3.0, 0.0, 0.0]), np.array([0.0, 0.4, 0.0])
    rB, vB = rA + np.array([-10.0, 0.05, 0.01]), vA + np.array([1.0, 0.0, 0.0])
    s = ics.System(1e-14, 1e-3, np.array([rA, rB]), np.array([vA, vB]), np.array([FLYBY_GM * 1.0000001, FLYBY_GM]),
                   np.array([1e-5, 1e-5]), np.zeros((0, 3)), np.zeros((0, 3)))
    o = O.OracleSystem(O.SYMBA, s, FLYBY_DT, lclose=False, rhill=np.array([0.5, 0.5]), L=L)
    st = o.state()
    mu_pair = float(s.Gmass.sum())
    r0 = st["pl_rh"][1] - st["pl_rh"][0]
    v0 = st["pl_vb"][1] - st["pl_vb"][0]
    E0, L0, A0 = _invariants(mu_pair, r0, v0)
    depth = -1
    for _ in range(FLYBY_STEPS):
        assert o.step() == 0
        depth = max(depth, o.stats()["max_depth"])
    st = o.state()
    r1 = st["pl_rh"][1] - st["pl_rh"][0]
    v1 = st["pl_vb"][1] - st["pl_vb"][0]
    E1, L1, A1 = _invariants(mu_pair, r1, v1)
    ra, _ = _hyperbola(mu_pair, r0, v0, FLYBY_STEPS * FLYBY_DT)
    return dict(dE=abs((E1 - E0) / E0), dL=float(np.linalg.norm(L1 - L0) / np.linalg.norm(L0)),
                dA=float(np.linalg.norm(A1 - A0) / np.linalg.norm(A0)),
                dr=float(np.linalg.norm(r1 - ra) / np.linalg.norm(ra)), depth=depth, sep0=float(np.linalg.norm(r0)),
                sep1=float(np.linalg.norm(r1)))


# Bounds about 100x what the oracle reaches (dE 3.8e-7, dA 5.5e-8, dr
# 1.6e-7: set by the second-order splitting of the step while the bodies
# are far apart, (dt omega)^2 ~ 2e-7 at 10 AU; dL by rounding).  Dropping
# the taper gives dE 7e-3, dropping the -1 kicks dA 0.14.
FLYBY_BOUNDS = dict(dE=4e-5, dL=1e-12, dA=5e-6, dr=2e-5)


def flyby_ok(res):
    return all(res[k] <= b for k, b in FLYBY_BOUNDS.items())


def test_hyperbola_helper_conserves_invariants():
    r0, v0 = np.array([-1.5, 0.05, 0.01]), np.array([1.0, 0.0, 0.0])
    r1, v1 = _hyperbola(0.08, r0, v0, 3.0)
    r2, v2 = _hyperbola(0.08, r1, -v1, 3.0)   # time reversal brings it back
    np.testing.assert_allclose(r2, r0, rtol=0, atol=1e-12)
    E0, L0, A0 = _invariants(0.08, r0, v0)
    E1, L1, A1 = _invariants(0.08, r1, v1)
    assert abs(E1 - E0) <= 1e-14 * abs(E0)
    np.testing.assert_allclose(L1, L0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(A1, A0, rtol=0, atol=1e-12)


def test_two_body_flyby_through_the_shells_matches_the_hyperbola():
    res = flyby()
    assert res["depth"] >= 7, res          # the pericentre lies eight shells down
    assert res["sep0"] > 9.9 and res["sep1"] > 9.9, res   # the bodies leave the shells again
    assert flyby_ok(res), res


# ------------------------------------------------------ planted encounters --
#
# Straight-line encounters: the central body and all bodies are nearly
# massless, so every pair's relative motion is r(t) = r0 + v t.  All Hill
# radii are h, so the pair radius of level d is R_d = 13 h RSHELL^d
# (symba_util.f90:245-267); pl-tp uses the massive body's alone, 6.5 h
# RSHELL^d.  The step is dt = 1 yr, relative speeds are 3 AU/yr.
#
# A pair whose closest approach b happens at t* is flagged at level d+1
# during a level-d substep when it approaches at the substep's start and its
# smallest separation over the substep is at most R_{d+1}
# (encounter_check.f90:573-621).  With t* in {1/6, 1/2, 5/6}, t* is the
# midpoint of the level-d substep that holds it, at every d, so the
# substep before it ends 3 * 3^-d / 2 AU short of closest approach, and the
# one after starts receding.  Choosing R_0 = 0.4 AU (R_1..R_5 = 0.1923,
# 0.0925, 0.0445, 0.0214, 0.0103):
#   level 1: substep before ends 0.5 AU out   > R_2
#   level 2:                     0.1667      > R_3
#   level 3:                     0.0556      > R_4
#   level 4:                     0.0185      > R_5
# so only the substep that holds t* descends, and a pair with
# R_{D+1} < b <= R_D reaches depth D.  Each level entered adds 3 substeps;
# level 0 has one.

H_PL = 0.4 / 13.0


def _planted(rel, GMs, tp=()):
    """Bodies at 3 AU from a central body of GM 1e-14 moving with a common
    velocity; rel[k] = (offset, relative velocity) of body k from body 0."""
    base_r, base_v = np.array([3.0, 0.0, 0.0]), np.array([0.0, 0.4, 0.0])
    r = np.array([base_r + np.asarray(o, float) for o, _ in rel])
    v = np.array([base_v + np.asarray(w, float) for _, w in rel])
    tr = np.array([base_r + np.asarray(o, float) for o, _ in tp]).reshape(-1, 3)
    tv = np.array([base_v + np.asarray(w, float) for _, w in tp]).reshape(-1, 3)
    s = ics.System(1e-14, 1e-3, r, v, np.asarray(GMs, float), np.full(len(GMs), 1e-9), tr, tv)
    return s


def _min_sep(o1, w1, o2, w2):
    dr = np.subtract(o2, o1)
    dv = np.subtract(w2, w1)
    t = -float(dr @ dv) / float(dv @ dv) if np.any(dv) else 0.0
    t = min(max(t, 0.0), 1.0)
    return float(np.linalg.norm(dr + dv * t)), t


CASES = {
    # A-B pass at t* = 1/2 with b = 0.015 (R_5 < b <= R_4): depth 4,
    # substeps 1 + 3 * 4
    "two_body": dict(
        rel=[((0, 0, 0), (0, 0, 0)), ((-1.5, 0.015, 0), (3, 0, 0))],
        GM=[3e-22, 2e-22], pairs=[(1, 2)], depth=4, nsub=13, levelm=[4, 4]),
    # chain B-A-C: A-C at t* = 1/6 with b = 0.06 (R_3 < b <= R_2, depth 2,
    # entered from level-1 substep 1), A-B at t* = 1/2 depth 4 (from
    # substep 2).  B and C never come within 0.70 AU.  Substeps: 1 + 3
    # (level 1) + 3 + 3 (level 2, twice) + 3 (level 3) + 3 (level 4) = 16
    "three_body_chain": dict(
        rel=[((0, 0, 0), (0, 0, 0)), ((-1.5, 0.015, 0), (3, 0, 0)), ((0, 0.06, 0.5), (0, 0, -3))],
        GM=[3e-22, 2e-22, 1e-22], pairs=[(1, 2), (1, 3)], depth=4, nsub=16, levelm=[4, 4, 2]),
    # chain A-B-C-D: A-B at 1/2 (b = 0.015, depth 4), B-C at 5/6 (b = 0.03,
    # R_4 < b <= R_3, depth 3), C-D at 1/6 (b = 0.06, depth 2); the other
    # pairs stay beyond 0.70 AU.  Level 2 is entered from all three level-1
    # substeps, level 3 twice (A-B, B-C), level 4 once:
    # 1 + 3 + 9 + 6 + 3 = 22
    "four_body_chain": dict(
        rel=[((-1.5, 0.015, 0), (3, 0, 0)), ((0, 0, 0), (0, 0, 0)), ((0, -2.5, 0.03), (0, 3, 0)),
             ((0.06, -2.5, -0.47), (0, 3, 3))],
        GM=[4e-22, 3e-22, 2e-22, 1e-22], pairs=[(1, 2), (2, 3), (3, 4)], depth=4, nsub=22, levelm=[4, 4, 3, 2]),
}

# pl-tp: a particle passes the massive body at t* = 1/2, b = 0.01, with
# pl-tp radii R_d = 6.5 h RSHELL^d for h = 0.2 / 6.5 (R_4 = 0.01068,
# R_5 = 0.00513; substeps before t* end 0.5, 0.1667, 0.0556, 0.0185 AU out,
# beyond R_2..R_5 = 0.0462, 0.0222, 0.0107, 0.0051): depth 4, 13 substeps.
TP_CASE = dict(rel=[((0, 0, 0), (0, 0, 0))], tp=[((-1.5, 0.01, 0), (3, 0, 0))], GM=[1e-22], h=0.2 / 6.5,
               depth=4, nsub=13, levelm=[4], tlevelm=[4])


def _premises(case):
    """Check the geometry the hand derivation assumes: every listed pair's
    closest approach is where the comment says, every other pair stays
    beyond 1.5 R_0."""
    rel = case["rel"]
    R0 = 13 * H_PL
    listed = {(i - 1, j - 1) for i, j in case["pairs"]}
    for i in range(len(rel)):
        for j in range(i + 1, len(rel)):
            d, t = _min_sep(rel[i][0], rel[i][1], rel[j][0], rel[j][1])
            if (i, j) in listed:
                assert d <= R0 and any(abs(t - x) < 1e-12 for x in (1 / 6, 1 / 2, 5 / 6)), (i, j, d, t)
            else:
                assert d > 1.5 * R0, (i, j, d)


def run_planted(name, L=None):
    case = CASES[name]
    _premises(case)
    s = _planted(case["rel"], case["GM"])
    o = O.OracleSystem(O.SYMBA, s, 1.0, lclose=False, rhill=np.full(len(case["GM"]), H_PL), L=L)
    assert o.step() == 0
    st, lv = o.stats(), o.levels()
    i1, i2, _ = o.enc_list(0)
    return dict(pairs=list(zip(i1.tolist(), i2.tolist())), depth=st["max_depth"], nsub=st["nsubsteps"],
                nenc=st["nplpl_enc"], levelm=lv["levelm"].tolist(), levelg=lv["levelg"].tolist())


@pytest.mark.parametrize("name", sorted(CASES))
def test_planted_encounters_levels_and_substeps(name):
    case = CASES[name]
    got = run_planted(name)
    assert got["pairs"] == case["pairs"]
    assert got["nenc"] == len(case["pairs"])
    assert got["depth"] == case["depth"]
    assert got["nsub"] == case["nsub"]
    assert got["levelm"] == case["levelm"]
    # after the step every level is back at 0 (set_recur_levels at each level's end)
    assert got["levelg"] == [0] * len(case["levelm"])


def test_planted_pltp_encounter_levels_and_substeps():
    c = TP_CASE
    d, t = _min_sep(c["rel"][0][0], c["rel"][0][1], c["tp"][0][0], c["tp"][0][1])
    assert d <= 6.5 * c["h"] and abs(t - 0.5) < 1e-12
    s = _planted(c["rel"], c["GM"], tp=c["tp"])
    o = O.OracleSystem(O.SYMBA, s, 1.0, lclose=False, rhill=np.full(1, c["h"]))
    assert o.step() == 0
    st, lv = o.stats(), o.levels()
    assert st["nplpl_enc"] == 0 and st["npltp_enc"] == 1
    assert st["max_depth"] == c["depth"] and st["nsubsteps"] == c["nsub"]
    assert lv["levelm"].tolist() == c["levelm"] and lv["tlevelm"].tolist() == c["tlevelm"]
    assert lv["levelg"].tolist() == [0] and lv["tlevelg"].tolist() == [0]


# ------------------------------------------------------------ whole disk --

def disk_energy_error(L=None):
    from test_oracle import _energy
    d = compact_disk(seed=9, npl=80, mhi=3e-6, alo=0.9, ahi=1.1)
    o = O.OracleSystem(O.SYMBA, d, 0.005, istep_out=20, L=L)
    st = o.state()
    Gm = d.Gmass[st["pl_id"] - 1]
    E0, L0 = _energy(d.GMcb, st["pl_rh"], st["pl_vh"], Gm)
    assert o.run(200) == 0
    st = o.state()
    E, L1 = _energy(d.GMcb, st["pl_rh"], st["pl_vh"], Gm)
    return abs((E - E0) / E0), abs((L1 - L0) / L0), o.stats()


# ------------------------------------------------------ planted mistakes --

@pytest.mark.parametrize("mutant", [1, 2])
def test_planted_mistakes_fail_a_pin(mutant, tmp_path):
    """SWO_MUTANT=1 drops the taper polynomial from the shell factor,
    SWO_MUTANT=2 leaves out the -1 shell kicks of every level below 0; each
    must fail at least one of the pins above, and the disk's energy bound
    of test_oracle.py."""
    L = O.load(O.build_variant(str(tmp_path), [f"SWO_MUTANT={mutant}"]))
    failed = []
    worst, wrong = shell_factor_errors(L, n=60)
    if wrong or worst >= 16.0:
        failed.append("shell factor")
    res = flyby(L)
    if not flyby_ok(res):
        failed.append(f"flyby {res}")
    dE, _, _ = disk_energy_error(L)
    if dE >= 3e-9:
        failed.append(f"disk energy {dE:.3g}")
    assert failed, "the planted mistake passed every pin"
    if mutant == 2:
        # the shell factor itself is right; the flyby must catch it
        assert any(f.startswith("flyby") for f in failed)
