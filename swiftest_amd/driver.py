"""The reference's driver loop on the resident path (swiftest_driver.f90:
130-180).

`run(integrator, param_path)` reads the parameter file and the ASCII initial
conditions (swiftest_amd.io), creates the resident system, and then, as the
reference's loop does: steps; evaluates the discard checks after every step
when any discard parameter is set (`nbody_system%discard`, :139); and at the
output cadence ISTEP_OUT (and after the last step) appends a frame to the
NetCDF output, with the energy record when ENERGY is set.  A frame at t0
precedes the loop.  With CHK_CLOSE, a SyMBA system without test particles
and COLLISION_MODEL MERGE (the reference's default), collisions are merged
inside the step (swg_system_set_merge; masses Gmass / GU) and the frames
follow the changing body set by id.  Other collisions stop the run with
SwiftestHipError SWG_ERR_COLLISION; their resolution is host work.
"""
from __future__ import annotations

import math
import os

from . import core, ics, io

_INTEGRATORS = {"WHM": core.WHM, "HELIO": core.HELIO, "SYMBA": core.SYMBA}


def run(integrator: str, param_path: str, out_path: str | None = None, nsteps: int | None = None,
        pair_mode: int = core.PAIR_EXACT):
    """Integrate the system of `param_path` with `integrator` (WHM, HELIO or
    SYMBA) from T0 to TSTOP (or for `nsteps` steps), writing frames to
    `out_path` (NetCDF) when given.  Returns the GpuSystem at the end."""
    prm = io.read_param(param_path)
    base = os.path.dirname(os.path.abspath(param_path))
    system, extra = io.read_in(prm, base)
    kind = _INTEGRATORS[integrator.upper()]
    dt, t0 = float(prm["DT"]), float(prm.get("T0", 0.0))
    if nsteps is None:
        nsteps = int(math.ceil((float(prm["TSTOP"]) - t0) / dt - 1e-9))
    istep_out = int(prm.get("ISTEP_OUT", -1))
    lclose = bool(prm.get("CHK_CLOSE", False))
    # G in the run's units when the parameter file gives them
    gu = io.gu(prm) if all(k in prm for k in ("MU2KG", "DU2M", "TU2S")) else None
    merge = (lclose and kind == core.SYMBA and system.ntp == 0 and
             str(prm.get("COLLISION_MODEL", "MERGE")).upper() == "MERGE")
    g = core.GpuSystem(kind, system, dt, lclose=lclose,
                       istep_out=max(istep_out, 0), rhill=extra.rhill, pair_mode=pair_mode,
                       gmtiny=max(float(prm.get("GMTINY", 0.0)), 0.0) if kind == core.SYMBA else 0.0,
                       merge=(gu if gu is not None else ics.GU, None) if merge else None)
    disc = dict(rmin=float(prm.get("CHK_RMIN", -1.0)), rmax=float(prm.get("CHK_RMAX", -1.0)),
                rmaxu=float(prm.get("CHK_EJECT", -1.0)), qmin=float(prm.get("CHK_QMIN", -1.0)))
    qr = prm.get("CHK_QMIN_RANGE", (-1.0, -1.0))
    disc["qmin_alo"], disc["qmin_ahi"] = float(qr[0]), float(qr[1])
    disc["qmin_coord"] = 1 if str(prm.get("CHK_QMIN_COORD", "HELIO")).upper() == "BARY" else 0
    discards = any(disc[k] >= 0.0 for k in ("rmin", "rmax", "rmaxu", "qmin"))
    energy = bool(prm.get("ENERGY", False))
    # room for the bodies merges can make: at most one per massive body
    writer = io.NetCDFWriter(out_path, system, extra, room=max(system.npl - 1, 0) if merge else 0) \
        if out_path else None

    def frame(t):
        if writer:
            writer.write_frame(t, g.state(), g.energy(GU=gu)[0] if energy else None,
                               g.bodies() if merge else None)

    frame(t0)
    k, iout = 0, 0
    try:
        while k < nsteps:
            # whole stretches up to the next frame in one call, unless the
            # discards are checked after every step
            m = 1 if discards else (min(istep_out - iout, nsteps - k) if istep_out > 0 else nsteps - k)
            g.run(m)
            k += m
            if discards:
                g.discard(**disc)
            if istep_out > 0:
                iout += m
                if iout == istep_out or k == nsteps:
                    iout = 0
                    frame(t0 + k * dt)
            elif k == nsteps:
                frame(t0 + k * dt)
    finally:
        if writer:
            writer.close()
    return g
