"""The reference's file formats: run parameters, ASCII initial conditions
and the NetCDF output schema.

- read_param / write_param cover the parameter file (swiftest_io_param_reader,
  swiftest_io.f90:2285-2628; swiftest_io_param_writer, :2800-2900).
- read_in / write_in cover the ASCII body files of IN_TYPE ASCII
  (swiftest_io_read_in_cb, :3163-3219; swiftest_io_read_frame_body,
  :3265-3380): the central body, then the massive bodies and the test
  particles, in XV or EL form.
- NetCDFWriter writes the output variables and dimensions the reference
  writes (netcdf_io_module.f90:20-144) with one record per output time.
  Python analysis code that reads the reference's output by name (xarray)
  can read it. The container is NetCDF-3 (scipy.io.netcdf_file), since no
  NetCDF-4/HDF5 library is installed.

Lines that start with '!' are comments, as in the reference's example files.
Fortran 'd' exponents are accepted.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import ics

GC = 6.6743e-11  # globals_module.f90:41

# keys whose values are numbers, flags (YES/T), or kept as upper-case strings
_REAL = {"T0", "TSTART", "TSTOP", "DT", "CHK_RMIN", "CHK_RMAX", "CHK_EJECT", "CHK_QMIN", "MU2KG", "TU2S", "DU2M",
         "GMTINY", "MIN_GMFRAG", "EORBIT_ORIG", "GMTOT_ORIG", "ECOLLISIONS", "EUNTRACKED"}
_INT = {"ISTEP_OUT", "NSTEP_OUT", "DUMP_CADENCE"}
_FLAG = {"CHK_CLOSE", "EXTRA_FORCE", "BIG_DISCARD", "RHILL_PRESENT", "ENERGY", "GR", "ROTATION", "TIDES",
         "FIRSTKICK", "FIRSTENERGY", "YORP", "YARKOVSKY"}
_UPPER = {"IN_TYPE", "IN_FORM", "OUT_TYPE", "OUT_FORM", "OUT_STAT", "CHK_QMIN_COORD", "INTERACTION_LOOPS",
          "ENCOUNTER_CHECK", "ENCOUNTER_CHECK_PLPL", "ENCOUNTER_CHECK_PLTP", "COLLISION_MODEL"}
_VEC = {"CHK_QMIN_RANGE": 2, "LTOT_ORIG": 3, "LORBIT_ORIG": 3, "LSPIN_ORIG": 3, "LESCAPE": 3}
# defaults of the fields the readers use (swiftest_module.f90 parameter type)
DEFAULTS = {"T0": 0.0, "DT": -1.0, "ISTEP_OUT": -1, "CHK_CLOSE": False, "CHK_RMIN": -1.0, "CHK_RMAX": -1.0,
            "CHK_EJECT": -1.0, "CHK_QMIN": -1.0, "CHK_QMIN_COORD": "HELIO", "CHK_QMIN_RANGE": (-1.0, -1.0),
            "RHILL_PRESENT": False, "ROTATION": False, "GR": False, "ENERGY": False, "IN_TYPE": "NETCDF_DOUBLE",
            "IN_FORM": "XV", "GMTINY": -1.0, "ENCOUNTER_CHECK": "TRIANGULAR", "INTERACTION_LOOPS": "TRIANGULAR"}


def _num(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def _tokens(line: str) -> list[str]:
    line = line.split("!", 1)[0]
    return line.split()


def read_param(path: str) -> dict:
    """The parameter file as a dict keyed by the reference's (upper-case)
    names, with DEFAULTS for the keys it leaves out."""
    out = dict(DEFAULTS)
    with open(path) as f:
        for line in f:
            tok = _tokens(line)
            if not tok:
                continue
            key, vals = tok[0].upper(), tok[1:]
            if not vals:
                continue
            if key in _REAL:
                out[key] = _num(vals[0])
            elif key in _INT:
                out[key] = int(_num(vals[0]))
            elif key in _FLAG:
                v = vals[0].upper()
                out[key] = v in ("YES", "T") if key not in ("FIRSTKICK", "FIRSTENERGY") else v not in ("NO", "F")
            elif key in _VEC:
                out[key] = tuple(_num(v) for v in vals[:_VEC[key]])
            elif key in _UPPER:
                out[key] = vals[0].upper()
            else:
                out[key] = vals[0] if len(vals) == 1 else " ".join(vals)
    if "ENCOUNTER_CHECK" in out and "ENCOUNTER_CHECK_PLPL" not in out:
        out["ENCOUNTER_CHECK_PLPL"] = out["ENCOUNTER_CHECK_PLTP"] = out["ENCOUNTER_CHECK"]
    return out


def gu(param: dict) -> float:
    """G in the run's units (swiftest_io.f90:2639)."""
    return GC / (param["DU2M"] ** 3 / (param["MU2KG"] * param["TU2S"] ** 2))


def write_param(path: str, param: dict):
    with open(path, "w") as f:
        for k, v in param.items():
            if isinstance(v, bool):
                v = "YES" if v else "NO"
            elif isinstance(v, tuple):
                v = " ".join(repr(float(x)) for x in v)
            elif isinstance(v, float):
                v = repr(v)
            f.write(f"{k:<17s}{v}\n")


@dataclass
class Bodies:
    """What the ASCII inputs hold beyond ics.System."""
    cb_name: str = "cb"
    j2rp2: float = 0.0
    j4rp4: float = 0.0
    pl_names: list = field(default_factory=list)
    tp_names: list = field(default_factory=list)
    rhill: np.ndarray | None = None


def _data_lines(path):
    with open(path) as f:
        return [t for t in (_tokens(ln) for ln in f) if t]


def _el_to_xv(mu, rows):
    a, e, inc, capom, omega, capm = (np.array([r[k] for r in rows]) for k in range(6))
    deg = np.pi / 180.0
    return ics.el2xv(mu, a, e, inc * deg, capom * deg, omega * deg, capm * deg)


def read_in(param: dict, base: str = ".") -> tuple[ics.System, Bodies]:
    """Read CB_IN, PL_IN and TP_IN (IN_TYPE ASCII) into an ics.System."""
    if param.get("IN_TYPE", "ASCII") != "ASCII":
        raise ValueError("only IN_TYPE ASCII is read here")
    form = param.get("IN_FORM", "XV")
    if form not in ("XV", "EL"):
        raise ValueError(f"{form} is not a recognized format code for input files")
    lclose, lrhill, lrot = param.get("CHK_CLOSE", False), param.get("RHILL_PRESENT", False), param.get("ROTATION", False)
    cb = _data_lines(os.path.join(base, param["CB_IN"]))
    extra = Bodies(cb_name=cb[0][0], j2rp2=_num(cb[3][0]), j4rp4=_num(cb[4][0]))
    GMcb, Rcb = _num(cb[1][0]), _num(cb[2][0])

    pl = _data_lines(os.path.join(base, param["PL_IN"]))
    n, k = int(pl[0][0]), 1
    Gm, rad, rh_, rows, names = np.zeros(n), np.zeros(n), np.zeros(n), [], []
    for i in range(n):
        head = pl[k]
        names.append(head[0])
        Gm[i] = _num(head[1])
        if lrhill:
            rh_[i] = _num(head[2])
        k += 1
        if lclose:
            rad[i] = _num(pl[k][0])
            k += 1
        rows.append([_num(t) for t in pl[k][:3]] + [_num(t) for t in pl[k + 1][:3]])
        k += 2 + (2 if lrot else 0)
    extra.pl_names, extra.rhill = names, (rh_ if lrhill else None)
    rows = np.array(rows).reshape(n, 6)
    if form == "XV":
        pr, pv = rows[:, :3].copy(), rows[:, 3:].copy()
    else:
        pr, pv = _el_to_xv(GMcb + Gm, rows)

    tp = _data_lines(os.path.join(base, param["TP_IN"])) if param.get("TP_IN") else [["0"]]
    m, k = int(tp[0][0]), 1
    trows, tnames = [], []
    for _ in range(m):
        tnames.append(tp[k][0])
        trows.append([_num(t) for t in tp[k + 1][:3]] + [_num(t) for t in tp[k + 2][:3]])
        k += 3
    extra.tp_names = tnames
    trows = np.array(trows).reshape(m, 6)
    if form == "XV":
        tr, tv = trows[:, :3].copy(), trows[:, 3:].copy()
    else:
        tr, tv = _el_to_xv(np.full(m, GMcb), trows) if m else (np.zeros((0, 3)), np.zeros((0, 3)))
    s = ics.System(GMcb=GMcb, Rcb=Rcb, pl_rh=pr, pl_vh=pv, Gmass=Gm, radius=rad, tp_rh=tr, tp_vh=tv,
                   meta={"param": param})
    return s, extra


def write_in(param: dict, system: ics.System, extra: Bodies | None = None, base: str = "."):
    """Write CB_IN, PL_IN and TP_IN in XV form."""
    extra = extra or Bodies()
    f17 = lambda x: repr(float(x))
    with open(os.path.join(base, param["CB_IN"]), "w") as f:
        for v in (extra.cb_name, f17(system.GMcb), f17(system.Rcb), f17(extra.j2rp2), f17(extra.j4rp4)):
            f.write(f"{v}\n")
    names = extra.pl_names or [f"pl{i + 1}" for i in range(system.npl)]
    with open(os.path.join(base, param["PL_IN"]), "w") as f:
        f.write(f"{system.npl}\n")
        for i in range(system.npl):
            head = f"{names[i]} {f17(system.Gmass[i])}"
            if param.get("RHILL_PRESENT", False):
                head += f" {f17(extra.rhill[i])}"
            f.write(head + "\n")
            if param.get("CHK_CLOSE", False):
                f.write(f17(system.radius[i]) + "\n")
            f.write(" ".join(f17(x) for x in system.pl_rh[i]) + "\n")
            f.write(" ".join(f17(x) for x in system.pl_vh[i]) + "\n")
    tnames = extra.tp_names or [f"tp{i + 1}" for i in range(system.ntp)]
    with open(os.path.join(base, param["TP_IN"]), "w") as f:
        f.write(f"{system.ntp}\n")
        for i in range(system.ntp):
            f.write(f"{tnames[i]}\n")
            f.write(" ".join(f17(x) for x in system.tp_rh[i]) + "\n")
            f.write(" ".join(f17(x) for x in system.tp_vh[i]) + "\n")


class NetCDFWriter:
    """Frames of one system in the reference's output variables.

    Dimensions: time (record), name (central body, then massive bodies, then
    test particles, then room for `room` bodies made by merges), space
    (x, y, z) and string32.  Row q holds the body of id q (the central body
    is id 0; new bodies take ids after the last test particle).  Per frame:
    rh, vh (time, name, space), Gmass, radius, rhill (time, name), status,
    npl, ntp, and when given the energy record KE_orb, KE_spin, PE, BE, TE,
    L_orbit, L_spin (the fields of swg_system_energy).  Rows of bodies not
    in the frame (merged away, or not made yet) hold NaN and status
    INACTIVE."""

    def __init__(self, path: str, system: ics.System, extra: Bodies | None = None, room: int = 0):
        from scipy.io import netcdf_file
        extra = extra or Bodies()
        self.npl, self.ntp, self.room = system.npl, system.ntp, int(room)
        n = 1 + self.npl + self.ntp + self.room
        self.f = netcdf_file(path, "w", version=2)
        f = self.f
        f.createDimension("time", None)
        f.createDimension("name", n)
        f.createDimension("space", 3)
        f.createDimension("string32", 32)
        f.createVariable("time", "d", ("time",))
        sp = f.createVariable("space", "c", ("space",))
        sp[:] = np.array(list("xyz"), "S1")
        # a merged body is named for its id (the reference appends _MERGE and
        # the id to the larger impactor's name, collision_resolve.f90:449-457)
        names = [extra.cb_name] + (extra.pl_names or [f"pl{i + 1}" for i in range(self.npl)]) + \
            (extra.tp_names or [f"tp{i + 1}" for i in range(self.ntp)]) + \
            [f"MERGE{1 + self.npl + self.ntp + q:07d}" for q in range(self.room)]
        nm = f.createVariable("name", "c", ("name", "string32"))
        nm[:] = np.array([list(s[:32].ljust(32)) for s in names], "S1")
        idv = f.createVariable("id", "i", ("name",))
        idv[:] = np.arange(n, dtype=np.int32)
        pt = f.createVariable("particle_type", "c", ("name", "string32"))
        kinds = ["Central Body"] + ["Massive Body"] * self.npl + ["Test Particle"] * self.ntp + \
            ["Massive Body"] * self.room
        pt[:] = np.array([list(s.ljust(32)) for s in kinds], "S1")
        for v in ("rh", "vh"):
            f.createVariable(v, "d", ("time", "name", "space"))
        for v in ("Gmass", "radius", "rhill"):
            f.createVariable(v, "d", ("time", "name"))
        f.createVariable("status", "i", ("time", "name"))
        f.createVariable("npl", "i", ("time",))
        f.createVariable("ntp", "i", ("time",))
        for v in ("KE_orb", "KE_spin", "PE", "BE", "TE", "GMescape"):
            f.createVariable(v, "d", ("time",))
        for v in ("L_orbit", "L_spin"):
            f.createVariable(v, "d", ("time", "space"))
        self.k = 0
        self.system = system

    def write_frame(self, t: float, state: dict, energy=None, bodies: dict | None = None):
        """One frame from GpuSystem.state() (one system) and, optionally, its
        energy row (GpuSystem.energy()[s]) and its bodies (GpuSystem.bodies():
        the Gmass and radius of each id, needed once merges change them)."""
        f, k, npl, ntp = self.f, self.k, self.npl, self.ntp
        s = self.system
        n = 1 + npl + ntp + self.room
        f.variables["time"][k] = t
        # rows by id: the device holds the massive bodies sorted (storage row
        # r is body pl_id[r])
        ids = np.asarray(state.get("pl_id", np.arange(1, npl + 1)))
        nan = np.nan

        def rows(a, fill, width=None):
            a = np.asarray(a)
            out = np.full((n,) + a.shape[1:], fill, dtype=a.dtype if a.dtype != np.float64 else np.float64)
            out[ids] = a
            return out

        rh, vh = rows(state["pl_rh"], nan), rows(state["pl_vh"], nan)
        rh[0] = vh[0] = 0.0
        rh[1 + npl:1 + npl + ntp], vh[1 + npl:1 + npl + ntp] = state["tp_rh"], state["tp_vh"]
        f.variables["rh"][k] = rh
        f.variables["vh"][k] = vh
        if bodies is None:
            gm = np.full(n, nan)
            rad = np.full(n, nan)
            gm[1:1 + npl], rad[1:1 + npl] = s.Gmass, s.radius
        else:
            gm, rad = rows(bodies["Gmass"], nan), rows(bodies["radius"], nan)
        gm[0], rad[0] = s.GMcb, s.Rcb
        gm[1 + npl:1 + npl + ntp] = rad[1 + npl:1 + npl + ntp] = 0.0
        f.variables["Gmass"][k] = gm
        f.variables["radius"][k] = rad
        rhill = rows(state.get("rhill", np.zeros(len(ids))), nan)
        rhill[0] = 0.0
        rhill[1 + npl:1 + npl + ntp] = 0.0
        f.variables["rhill"][k] = rhill
        status = rows(np.asarray(state["pl_status"], np.int32), np.int32(1))
        status[0] = 0
        status[1 + npl:1 + npl + ntp] = state["tp_status"]
        f.variables["status"][k] = status.astype(np.int32)
        f.variables["npl"][k] = int((state["pl_status"] == 0).sum())
        f.variables["ntp"][k] = int((state["tp_status"] == 0).sum())
        # every record variable gets a value in every frame (the record
        # layout needs them all); without an energy record those are NaN
        e = np.asarray(energy) if energy is not None else np.full(16, np.nan)
        for name, j in (("KE_orb", 0), ("KE_spin", 1), ("PE", 2), ("BE", 3), ("TE", 5)):
            f.variables[name][k] = e[j]
        f.variables["L_orbit"][k] = e[6:9]
        f.variables["L_spin"][k] = e[9:12]
        f.variables["GMescape"][k] = 0.0 if energy is not None else np.nan
        self.k += 1

    def close(self):
        self.f.close()
