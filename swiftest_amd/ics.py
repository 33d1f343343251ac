"""Offline synthetic initial conditions for the benchmark workloads.

Units are those of the reference examples: M_sun, AU, yr.  The gravitational
constant in these units is built the way the reference builds it
(src/swiftest/swiftest_io.f90:2639: GU = GC / (DU^3 / (MU TU^2))) from the
IAU nominal GM_sun the reference's Python front end takes from astropy
(swiftest/constants.py:17-27).

Random draws use numpy's PCG64 with a fixed seed per workload, so a workload
is fully determined by (name, seed).  Orbital elements are converted to
heliocentric Cartesian state with mu = GM_cb + Gm for massive bodies and
mu = GM_cb for test particles, as the reference does after set_mu
(swiftest_util.f90:2439-2443).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GC = 6.6743e-11                   # globals_module.f90:41
AU2M = 1.495978707e11
GMSUN_SI = 1.3271244e20
MSUN = GMSUN_SI / GC
YR2S = 365.25 * 86400.0
GU = GC / (AU2M ** 3 / (MSUN * YR2S ** 2))   # G in AU^3 / (M_sun yr^2)
GMSUN = GU * 1.0                                # ~39.4769264
RHO_3GCC = 3000.0 * AU2M ** 3 / MSUN            # 3 g/cm^3 in M_sun / AU^3
RSUN_AU = 6.957e8 / AU2M
DEG = np.pi / 180.0


@dataclass
class System:
    """One planetary system: central body plus massive bodies and test particles."""
    GMcb: float
    Rcb: float
    pl_rh: np.ndarray          # (npl, 3)
    pl_vh: np.ndarray          # (npl, 3)
    Gmass: np.ndarray          # (npl,)
    radius: np.ndarray         # (npl,)
    tp_rh: np.ndarray          # (ntp, 3)
    tp_vh: np.ndarray          # (ntp, 3)
    meta: dict = field(default_factory=dict)

    @property
    def npl(self) -> int:
        return int(self.Gmass.shape[0])

    @property
    def ntp(self) -> int:
        return int(self.tp_rh.shape[0])


def kepler_E(M: np.ndarray, e: np.ndarray) -> np.ndarray:
    """Eccentric anomaly from mean anomaly (elliptic), Newton iteration."""
    E = M + e * np.sin(M)
    if E.size == 0:
        return E
    for _ in range(50):
        f = E - e * np.sin(E) - M
        E_new = E - f / (1.0 - e * np.cos(E))
        if np.max(np.abs(E_new - E)) < 1e-15:
            E = E_new
            break
        E = E_new
    return E


def el2xv(mu, a, e, inc, capom, omega, capm):
    """Elliptic orbital elements (radians) to heliocentric position/velocity."""
    mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), np.shape(a))
    E = kepler_E(np.mod(capm, 2 * np.pi), e)
    cE, sE = np.cos(E), np.sin(E)
    sq = np.sqrt(1.0 - e * e)
    n = np.sqrt(mu / a ** 3)
    xo = a * (cE - e)
    yo = a * sq * sE
    den = 1.0 - e * cE
    vxo = -a * n * sE / den
    vyo = a * n * sq * cE / den
    cO, sO = np.cos(capom), np.sin(capom)
    cw, sw = np.cos(omega), np.sin(omega)
    ci, si = np.cos(inc), np.sin(inc)
    p = np.stack([cO * cw - sO * sw * ci, sO * cw + cO * sw * ci, sw * si], axis=-1)
    q = np.stack([-cO * sw - sO * cw * ci, -sO * sw + cO * cw * ci, cw * si], axis=-1)
    r = xo[:, None] * p + yo[:, None] * q
    v = vxo[:, None] * p + vyo[:, None] * q
    return np.ascontiguousarray(r), np.ascontiguousarray(v)


def _loguniform(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def _angles(rng, n):
    return (rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n))


def _radius_from_mass(m):
    return (3.0 * m / (4.0 * np.pi * RHO_3GCC)) ** (1.0 / 3.0)


def _planets(rng, m, a, e, inc, GMcb):
    capom, omega, capm = _angles(rng, m.size)
    Gm = GU * m
    r, v = el2xv(GMcb + Gm, a, e, inc, capom, omega, capm)
    return r, v, Gm, _radius_from_mass(m)


def _particles(rng, a, e, inc, GMcb):
    capom, omega, capm = _angles(rng, a.size)
    return el2xv(GMcb, a, e, inc, capom, omega, capm)


def whm_solar(seed: int = 1, npl: int = 8, ntp: int = 1000) -> System:
    """Central body of 1 M_sun, npl planets (m log-uniform 1e-7..1e-3 M_sun,
    a log-uniform 0.4..30 AU, e U(0,0.05), i U(0,2 deg)) and ntp test particles
    (a U(0.5,40), e U(0,0.1), i U(0,5 deg))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = _loguniform(rng, 1e-7, 1e-3, npl)
    a = _loguniform(rng, 0.4, 30.0, npl)
    e = rng.uniform(0, 0.05, npl)
    i = rng.uniform(0, 2.0, npl) * DEG
    r, v, Gm, R = _planets(rng, m, a, e, i, GMSUN)
    at = rng.uniform(0.5, 40.0, ntp)
    et = rng.uniform(0, 0.1, ntp)
    it = rng.uniform(0, 5.0, ntp) * DEG
    tr, tv = _particles(rng, at, et, it, GMSUN)
    return System(GMSUN, RSUN_AU, r, v, Gm, R, tr, tv,
                  dict(name="whm_8pl_1000tp", seed=seed, dt=0.01, nsteps=10000))


def symba_disk(seed: int = 2, npl: int = 2000) -> System:
    """Planetesimal disk: dN/dm ~ m^-2 on 1e-9..1e-7 M_sun, radius from
    3 g/cm^3, a U(0.7,1.5) AU, e and i Rayleigh(0.01) (i in radians)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(0, 1, npl)
    m1, m2 = 1e-9, 1e-7
    m = 1.0 / (1.0 / m1 - u * (1.0 / m1 - 1.0 / m2))
    a = rng.uniform(0.7, 1.5, npl)
    e = np.minimum(0.01 * np.sqrt(-2.0 * np.log1p(-rng.uniform(0, 1, npl))), 0.9)
    i = 0.01 * np.sqrt(-2.0 * np.log1p(-rng.uniform(0, 1, npl)))
    r, v, Gm, R = _planets(rng, m, a, e, i, GMSUN)
    z = np.zeros((0, 3))
    return System(GMSUN, RSUN_AU, r, v, Gm, R, z, z.copy(),
                  dict(name="symba_disk_2000pl", seed=seed, dt=0.005, nsteps=1000))


def symba_big(seed: int = 3, npl: int = 32768, ntp: int = 1048576) -> System:
    """Large SyMBA system: npl bodies with the whm_solar mass/radius/orbit laws
    over a U(0.5,5) AU, plus ntp test particles with a U(0.3,10) AU,
    e U(0,0.2), i U(0,10 deg)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = _loguniform(rng, 1e-7, 1e-3, npl)
    a = rng.uniform(0.5, 5.0, npl)
    e = rng.uniform(0, 0.05, npl)
    i = rng.uniform(0, 2.0, npl) * DEG
    r, v, Gm, R = _planets(rng, m, a, e, i, GMSUN)
    at = rng.uniform(0.3, 10.0, ntp)
    et = rng.uniform(0, 0.2, ntp)
    it = rng.uniform(0, 10.0, ntp) * DEG
    tr, tv = _particles(rng, at, et, it, GMSUN)
    return System(GMSUN, RSUN_AU, r, v, Gm, R, tr, tv,
                  dict(name=f"symba_{npl}pl_{ntp}tp", seed=seed, dt=0.005, nsteps=200))


def ensemble(seed: int = 4, nsys: int = 4096, nmin: int = 4, nmax: int = 512, tpmax: int = 2000) -> list[System]:
    """Independent systems: N log-uniform over nmin..nmax, m log-uniform
    1e-8..1e-4 M_sun, a log-uniform 0.3..30 AU, e U(0,0.1), i U(0,3 deg), and
    U{0..tpmax} test particles drawn from the same orbit laws."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nsys):
        n = int(np.floor(_loguniform(rng, nmin, nmax + 1, 1)[0]))
        n = max(nmin, min(nmax, n))
        m = _loguniform(rng, 1e-8, 1e-4, n)
        a = _loguniform(rng, 0.3, 30.0, n)
        e = rng.uniform(0, 0.1, n)
        i = rng.uniform(0, 3.0, n) * DEG
        r, v, Gm, R = _planets(rng, m, a, e, i, GMSUN)
        nt = int(rng.integers(0, tpmax + 1))
        at = _loguniform(rng, 0.3, 30.0, nt)
        et = rng.uniform(0, 0.1, nt)
        it = rng.uniform(0, 3.0, nt) * DEG
        tr, tv = _particles(rng, at, et, it, GMSUN) if nt else (np.zeros((0, 3)), np.zeros((0, 3)))
        out.append(System(GMSUN, RSUN_AU, r, v, Gm, R, tr, tv, dict(dt=0.01)))
    return out
