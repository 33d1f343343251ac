"""Shared builders for the parity tests."""
from __future__ import annotations

import numpy as np

from swiftest_amd import ics


def compact_disk(seed=11, npl=160, ntp=0, mlo=1e-7, mhi=1e-5, alo=0.95, ahi=1.05, sig=0.01):
    """A crowded SyMBA disk whose bodies are in close encounters from step 0,
    so that the recursion reaches several levels within a few steps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = np.exp(rng.uniform(np.log(mlo), np.log(mhi), npl))
    a = rng.uniform(alo, ahi, npl)
    e = sig * np.sqrt(-2.0 * np.log1p(-rng.uniform(0, 1, npl)))
    i = sig * np.sqrt(-2.0 * np.log1p(-rng.uniform(0, 1, npl)))
    capom, omega, capm = (rng.uniform(0, 2 * np.pi, npl) for _ in range(3))
    Gm = ics.GU * m
    r, v = ics.el2xv(ics.GMSUN + Gm, a, e, i, capom, omega, capm)
    R = ics._radius_from_mass(m)
    if ntp:
        at = rng.uniform(alo, ahi, ntp)
        et = rng.uniform(0, 0.02, ntp)
        it = rng.uniform(0, 0.02, ntp)
        tr, tv = ics.el2xv(ics.GMSUN, at, et, it, *(rng.uniform(0, 2 * np.pi, ntp) for _ in range(3)))
    else:
        tr, tv = np.zeros((0, 3)), np.zeros((0, 3))
    return ics.System(ics.GMSUN, ics.RSUN_AU, r, v, Gm, R, tr, tv, dict(name="compact_disk", seed=seed))


def sparse_system(seed=21, npl=12, ntp=300):
    """A planetary system with few encounters and many test particles."""
    return ics.whm_solar(seed=seed, npl=npl, ntp=ntp)


def ulp_diff(a, b):
    """Largest distance in units in the last place between two float64 arrays."""
    a = np.ascontiguousarray(a, np.float64).view(np.int64)
    b = np.ascontiguousarray(b, np.float64).view(np.int64)
    a = np.where(a < 0, np.int64(-0x8000000000000000) - a, a)
    b = np.where(b < 0, np.int64(-0x8000000000000000) - b, b)
    return int(np.max(np.abs(a - b))) if a.size else 0
