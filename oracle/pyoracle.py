"""ctypes binding of the CPU restatement (libswiftest_oracle*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict = {}

D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("lclose", C.c_int), ("short_circuit", C.c_int),
                ("istep_out", C.c_int), ("rhill_present", C.c_int), ("nthreads", C.c_int),
                ("gmtiny", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("nplpl_enc", C.c_int64), ("npltp_enc", C.c_int64), ("max_depth", C.c_int32),
                ("nsubsteps", C.c_int64), ("ncollision", C.c_int64), ("ndrifterr", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(omp: bool = False):
    key = "omp" if omp else "strict"
    if key not in _LIBS:
        name = "libswiftest_oracle_omp.so" if omp else "libswiftest_oracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        _LIBS[key] = load(path)
    return _LIBS[key]


def build_variant(outdir, defines=()):
    """Compile swiftest_oracle.c with extra -D macros into outdir (the
    planted-mistake builds of the pin tests); returns the library path."""
    out = os.path.join(outdir, "libswiftest_oracle_" + "_".join(d.replace("=", "") for d in defines) + ".so")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v3", "-fPIC",
                    "-shared", "-w", *[f"-D{d}" for d in defines], "-o", out,
                    os.path.join(HERE, "swiftest_oracle.c"), "-lm"], check=True)
    return out


def load(path):
    """Open an oracle build and declare its entry points."""
    L = C.CDLL(path)
    vp = C.c_void_p
    L.swo_kick_getacch_int_all_tri.argtypes = [C.c_int, C.c_int, D, D, vp, D]
    L.swo_kick_getacch_int_all_flat.argtypes = [C.c_int, C.c_int64, I32, D, D, vp, D]
    L.swo_kick_getacch_int_all_tp.argtypes = [C.c_int, C.c_int, D, D, D, I32, D]
    L.swo_drift_all.argtypes = [D, D, D, C.c_int, C.c_double, I32, I32, C.c_int, C.c_double]
    L.swo_drift_kepu_stumpff.argtypes = [C.c_double] + [C.POINTER(C.c_double)] * 4
    L.swo_encounter_check_one.argtypes = [C.c_double] * 8 + [C.POINTER(C.c_int32)] * 2
    pp = C.POINTER(C.POINTER(C.c_int32))
    L.swo_encounter_check_all_plpl.argtypes = [C.c_int, D, D, D, C.c_double, pp, pp, pp]
    L.swo_encounter_check_all_plpl.restype = C.c_int64
    L.swo_encounter_check_all_plplm.argtypes = [C.c_int, C.c_int, D, D, D, C.c_double, pp, pp, pp]
    L.swo_encounter_check_all_plplm.restype = C.c_int64
    L.swo_encounter_check_all_pltp.argtypes = [C.c_int, C.c_int, D, D, D, D, D, C.c_double, pp, pp, pp]
    L.swo_encounter_check_all_pltp.restype = C.c_int64
    L.swo_free.argtypes = [vp]
    L.swo_collision_check_one.argtypes = [C.c_double] * 9 + [C.c_int32] + [C.POINTER(C.c_int32)] * 2
    L.swo_orbel_xv2aeq.argtypes = [C.c_double] * 7 + [C.POINTER(C.c_double)] * 3
    L.swo_orbel_xv2a.argtypes = [C.c_double] * 7
    L.swo_orbel_xv2a.restype = C.c_double
    L.swo_pow_m1p5.argtypes = [C.c_double]
    L.swo_pow_m1p5.restype = C.c_double
    L.swo_powi.argtypes = [C.c_double, C.c_int]
    L.swo_powi.restype = C.c_double
    L.swo_sys_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, D, D, D, vp, D, D,
                              C.POINTER(Params)]
    L.swo_sys_new.restype = vp
    L.swo_sys_free.argtypes = [vp]
    L.swo_sys_run.argtypes = [vp, C.c_int64]
    L.swo_sys_step.argtypes = [vp]
    L.swo_sys_get.argtypes = [vp] + [vp] * 8
    L.swo_sys_get_misc.argtypes = [vp] * 6
    L.swo_sys_get_order.argtypes = [vp, I32]
    L.swo_sys_stats.argtypes = [vp, C.POINTER(Stats)]
    L.swo_sys_npl.argtypes = [vp]
    L.swo_sys_ntp.argtypes = [vp]
    L.swo_sys_list.argtypes = [vp, C.c_int, vp, vp, vp]
    L.swo_sys_list.restype = C.c_int64
    L.swo_sys_list_status.argtypes = [vp, C.c_int, vp]
    L.swo_sys_list_status.restype = C.c_int64
    L.swo_sys_set_collision_mode.argtypes = [vp, C.c_int]
    L.swo_sys_set_collision_mode.restype = None
    L.swo_sys_set_merge.argtypes = [vp, C.c_double, vp]
    L.swo_sys_get_bodies.argtypes = [vp] * 5
    L.swo_sys_get_bodies.restype = None
    L.swo_sys_nmerge.argtypes = [vp]
    L.swo_sys_nmerge.restype = C.c_int64
    L.swo_norm2.argtypes = [C.c_double] * 3
    L.swo_norm2.restype = C.c_double
    L.swo_coord_vh2vb.argtypes = [C.c_int, C.c_double, D, D, D, D]
    L.swo_sys_discard.argtypes = [vp, D, C.c_int, C.POINTER(C.c_int64)]
    L.swo_discard.argtypes = ([C.c_int] * 3 + [C.c_double] * 2 + [D, D, D, D, D, D, D, I32, I32, I32, D, D, I32]
                              + [D, D, D, I32, I32, D, D, D, C.c_int, C.c_int, C.POINTER(C.c_int64)])
    L.swo_energy.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, D, D, D, D, D, D, I32, C.c_int, D, vp]
    L.swo_sys_get_levels.argtypes = [vp] * 5
    L.swo_shell_fac.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.swo_shell_fac.restype = C.c_int
    return L


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- leaf kernels -------------------------------------------------------

def accel_plpl(r, Gm, radius=None, acc=None, nplm=None, omp=False):
    r = _c(r)
    n = r.shape[0]
    acc = np.zeros((n, 3)) if acc is None else _c(acc).copy()
    rad = None if radius is None else _c(radius)
    lib(omp).swo_kick_getacch_int_all_tri(n, n if nplm is None else nplm, r, _c(Gm), _p(rad), acc)
    return acc


def accel_pairs(k_plpl, r, Gm, radius=None, acc=None):
    r = _c(r)
    n = r.shape[0]
    k = _c(k_plpl, np.int32).reshape(-1, 2)
    acc = np.zeros((n, 3)) if acc is None else _c(acc).copy()
    rad = None if radius is None else _c(radius)
    lib().swo_kick_getacch_int_all_flat(n, k.shape[0], k, r, _c(Gm), _p(rad), acc)
    return acc


def accel_pltp(rtp, rpl, GMpl, lmask=None, acc=None, omp=False):
    rtp = _c(rtp)
    ntp = rtp.shape[0]
    rpl = _c(rpl)
    lm = np.ones(ntp, np.int32) if lmask is None else _c(lmask, np.int32)
    acc = np.zeros((ntp, 3)) if acc is None else _c(acc).copy()
    lib(omp).swo_kick_getacch_int_all_tp(ntp, rpl.shape[0], rtp, rpl, _c(GMpl), lm, acc)
    return acc


def drift(mu, x, v, dt, lmask=None, omp=False):
    x = _c(x).copy()
    v = _c(v).copy()
    n = x.shape[0]
    mu = np.broadcast_to(np.asarray(mu, np.float64), (n,)).copy()
    lm = np.ones(n, np.int32) if lmask is None else _c(lmask, np.int32)
    fl = np.zeros(n, np.int32)
    lib(omp).swo_drift_all(mu, x, v, n, dt, lm, fl, 0, 0.0)
    return x, v, fl


def stumpff(x):
    c = [C.c_double() for _ in range(4)]
    lib().swo_drift_kepu_stumpff(x, *[C.byref(q) for q in c])
    return tuple(q.value for q in c)


def encounter_one(xr, vr, renc, dt):
    le, lv = C.c_int32(), C.c_int32()
    lib().swo_encounter_check_one(*map(float, xr), *map(float, vr), float(renc), float(dt), C.byref(le), C.byref(lv))
    return bool(le.value), bool(lv.value)


def _take_list(n, p1, p2, p3):
    L = lib()
    out = []
    for p in (p1, p2, p3):
        arr = np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].copy() if n else np.zeros(0, np.int32)
        L.swo_free(C.cast(p, C.c_void_p))
        out.append(arr)
    return out


def encounters_plpl(r, v, renc, dt, omp=False):
    p = [C.POINTER(C.c_int32)() for _ in range(3)]
    n = lib(omp).swo_encounter_check_all_plpl(len(renc), _c(r), _c(v), _c(renc), dt, *[C.byref(q) for q in p])
    return _take_list(n, *p)


def encounters_plplm(r, v, renc, nplm, dt, omp=False):
    p = [C.POINTER(C.c_int32)() for _ in range(3)]
    n = lib(omp).swo_encounter_check_all_plplm(len(renc), int(nplm), _c(r), _c(v), _c(renc), dt,
                                               *[C.byref(q) for q in p])
    return _take_list(n, *p)


def encounters_pltp(rpl, vpl, rtp, vtp, renc, dt, omp=False):
    p = [C.POINTER(C.c_int32)() for _ in range(3)]
    n = lib(omp).swo_encounter_check_all_pltp(len(renc), _c(rtp).shape[0], _c(rpl), _c(vpl), _c(rtp), _c(vtp),
                                              _c(renc), dt, *[C.byref(q) for q in p])
    return _take_list(n, *p)


def collision_one(xr, vr, Gmtot, rlim, dt, lvdotr):
    lc, lcl = C.c_int32(), C.c_int32()
    lib().swo_collision_check_one(*map(float, xr), *map(float, vr), Gmtot, rlim, dt, int(lvdotr),
                                  C.byref(lc), C.byref(lcl))
    return bool(lc.value), bool(lcl.value)


def xv2aeq(mu, r, v):
    a, e, q = C.c_double(), C.c_double(), C.c_double()
    lib().swo_orbel_xv2aeq(mu, *map(float, r), *map(float, v), C.byref(a), C.byref(e), C.byref(q))
    return a.value, e.value, q.value


def pow_m1p5(x):
    return lib().swo_pow_m1p5(float(x))


def powi(x, n):
    return lib().swo_powi(float(x), int(n))


def shell_fac(ri, r2, L=None):
    """(kept, fac) of the SyMBA shell kick for shell radius^2 ri and
    separation^2 r2 (symba_kick.f90:187-208)."""
    f = C.c_double(0.0)
    k = (L or lib()).swo_shell_fac(float(ri), float(r2), C.byref(f))
    return bool(k), f.value


# ---- whole systems --------------------------------------------------------

WHM, HELIO, SYMBA = 0, 1, 2


class OracleSystem:
    """A system integrated by the CPU restatement of the reference step layer."""

    def __init__(self, integrator, sysic, dt, lclose=True, short_circuit=True, istep_out=0,
                 rhill=None, nthreads=1, omp=False, L=None, gmtiny=0.0, collision_finish_step=False,
                 merge=None):
        self.L = L if L is not None else lib(omp)
        prm = Params(dt, int(lclose), int(short_circuit), int(istep_out), int(rhill is not None), int(nthreads),
                     float(gmtiny))
        self.npl, self.ntp = sysic.npl, sysic.ntp
        rh_in = None if rhill is None else _c(rhill)
        self._keep = rh_in
        self.h = self.L.swo_sys_new(integrator, self.npl, self.ntp, sysic.GMcb, sysic.Rcb, _c(sysic.pl_rh),
                                    _c(sysic.pl_vh), _c(sysic.Gmass), _c(sysic.radius), _p(rh_in),
                                    _c(sysic.tp_rh).reshape(-1, 3), _c(sysic.tp_vh).reshape(-1, 3), C.byref(prm))
        # a detected collision: stop at once (default), or flag the pairs and
        # finish the step (swo_sys_set_collision_mode)
        self.L.swo_sys_set_collision_mode(self.h, int(collision_finish_step))
        # merge = (GU, pl_mass or None): collisions resolved by merging
        # inside the step (swo_sys_set_merge)
        if merge is not None:
            GU, mass = merge
            self._mass = None if mass is None else _c(mass)
            if self.L.swo_sys_set_merge(self.h, float(GU), _p(self._mass)) != 0:
                raise ValueError("merging needs SyMBA, no test particles and GU > 0")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.swo_sys_free(self.h)
            self.h = None

    def run(self, nsteps):
        return self.L.swo_sys_run(self.h, int(nsteps))

    def step(self):
        return self.L.swo_sys_step(self.h)

    def discard(self, rmin=-1.0, rmax=-1.0, rmaxu=-1.0, qmin=-1.0, qmin_alo=-1.0, qmin_ahi=-1.0, qmin_coord=0):
        """swo_sys_discard: (massive bodies, test particles) newly discarded."""
        prm = np.array([rmin, rmax, rmaxu, qmin, qmin_alo, qmin_ahi], np.float64)
        counts = (C.c_int64 * 2)()
        self.L.swo_sys_discard(self.h, prm, int(qmin_coord), counts)
        return counts[0], counts[1]

    def state(self):
        self.npl = self.L.swo_sys_npl(self.h)
        npl, ntp = self.npl, self.ntp
        out = dict(pl_rh=np.zeros((npl, 3)), pl_vh=np.zeros((npl, 3)), pl_vb=np.zeros((npl, 3)),
                   tp_rh=np.zeros((ntp, 3)), tp_vh=np.zeros((ntp, 3)), tp_vb=np.zeros((ntp, 3)),
                   pl_status=np.zeros(npl, np.int32), tp_status=np.zeros(ntp, np.int32))
        self.L.swo_sys_get(self.h, *[_p(out[k]) for k in ("pl_rh", "pl_vh", "pl_vb", "tp_rh", "tp_vh", "tp_vb",
                                                           "pl_status", "tp_status")])
        rhill = np.zeros(npl)
        cb_vb, ptb, pte = np.zeros(3), np.zeros(3), np.zeros(3)
        lf = np.zeros(1, np.int32)
        self.L.swo_sys_get_misc(self.h, _p(rhill), _p(cb_vb), _p(ptb), _p(pte), _p(lf))
        ids = np.zeros(npl, np.int32)
        self.L.swo_sys_get_order(self.h, ids)
        out.update(rhill=rhill, cb_vb=cb_vb, ptbeg=ptb, ptend=pte, lfirst=int(lf[0]), pl_id=ids)
        return out

    def bodies(self):
        """Ids, masses, Gmass and radii of the massive bodies, in storage order."""
        n = self.L.swo_sys_npl(self.h)
        out = dict(id=np.zeros(n, np.int32), mass=np.zeros(n), Gmass=np.zeros(n), radius=np.zeros(n))
        self.L.swo_sys_get_bodies(self.h, *[_p(out[k]) for k in ("id", "mass", "Gmass", "radius")])
        return out

    @property
    def nmerge(self):
        return self.L.swo_sys_nmerge(self.h)

    def stats(self):
        st = Stats()
        self.L.swo_sys_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def levels(self):
        """levelg, levelm of the massive bodies and of the test particles
        after the last step, in storage order."""
        self.npl = self.L.swo_sys_npl(self.h)
        out = [np.zeros(self.npl, np.int32), np.zeros(self.npl, np.int32), np.zeros(self.ntp, np.int32),
               np.zeros(self.ntp, np.int32)]
        self.L.swo_sys_get_levels(self.h, *[_p(a) for a in out])
        return dict(levelg=out[0], levelm=out[1], tlevelg=out[2], tlevelm=out[3])

    def enc_list(self, which=0):
        n = self.L.swo_sys_list(self.h, which, None, None, None)
        a, b, c = (np.zeros(n, np.int32) for _ in range(3))
        self.L.swo_sys_list(self.h, which, _p(a), _p(b), _p(c))
        return a, b, c

    def enc_status(self, which=0):
        """Status of each pair of the level-0 list (0 ACTIVE, -14 COLLIDED)."""
        n = self.L.swo_sys_list_status(self.h, which, None)
        st = np.zeros(n, np.int32)
        self.L.swo_sys_list_status(self.h, which, _p(st))
        return st


# ---- energy and angular momentum ---------------------------------------

ENERGY_NOUT = 16


def norm2(x, y, z):
    return lib().swo_norm2(float(x), float(y), float(z))


def energy(GMcb, mcb, Rcb, cb_vb, rh, vb, Gmass, mass, radius=None, status=None, lclose=False):
    """swo_energy: the 16-field energy record of one system (and the
    barycentric positions)."""
    rh = _c(rh).reshape(-1, 3)
    n = rh.shape[0]
    st = np.zeros(n, np.int32) if status is None else _c(status, np.int32)
    rad = np.ones(n) if radius is None else _c(radius)
    out = np.zeros(ENERGY_NOUT)
    rb = np.zeros((max(n, 1), 3))
    lib().swo_energy(n, float(GMcb), float(mcb), float(Rcb), _c(cb_vb), rh, _c(vb).reshape(-1, 3), _c(Gmass),
                     _c(mass), rad, st, int(bool(lclose)), out, _p(rb))
    return out, rb[:n]


def vh2vb(GMcb, Gmass, vh):
    """swo_coord_vh2vb: (vb, cb_vb)."""
    vh = _c(vh).reshape(-1, 3)
    vb, cbv = np.zeros_like(vh), np.zeros(3)
    lib().swo_coord_vh2vb(vh.shape[0], float(GMcb), _c(Gmass), vh, vb, cbv)
    return vb, cbv


class DiscardState:
    """Arrays of one system for swo_discard (the state the checks read and
    change, pericentre-passage state included)."""

    def __init__(self, integrator, GMcb, Rcb, cb_vb, pl_rh, pl_vh, pl_vb, Gmass, radius, rhill, tp_rh, tp_vh, tp_vb,
                 pl_status=None, tp_status=None, nplenc=None):
        self.integ, self.GMcb, self.Rcb = int(integrator), float(GMcb), float(Rcb)
        self.cb_vb = _c(cb_vb).copy()
        self.pl_rh, self.pl_vh, self.pl_vb = (_c(a).reshape(-1, 3).copy() for a in (pl_rh, pl_vh, pl_vb))
        self.Gmass, self.radius, self.rhill = _c(Gmass).copy(), _c(radius).copy(), _c(rhill).copy()
        self.tp_rh, self.tp_vh, self.tp_vb = (_c(a).reshape(-1, 3).copy() for a in (tp_rh, tp_vh, tp_vb))
        n, m = self.pl_rh.shape[0], self.tp_rh.shape[0]
        self.pl_status = np.zeros(n, np.int32) if pl_status is None else _c(pl_status, np.int32).copy()
        self.tp_status = np.zeros(m, np.int32) if tp_status is None else _c(tp_status, np.int32).copy()
        self.nplenc = np.zeros(n, np.int32) if nplenc is None else _c(nplenc, np.int32).copy()
        self.pl_isperi, self.tp_isperi = np.ones(max(n, 1), np.int32), np.ones(max(m, 1), np.int32)
        self.pl_atp, self.pl_peri = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        self.tp_atp, self.tp_peri = np.zeros(max(m, 1)), np.zeros(max(m, 1))
        self.lfirst_peri = np.ones(1, np.int32)

    def discard(self, dt, rmin=-1.0, rmax=-1.0, rmaxu=-1.0, qmin=-1.0, qmin_alo=-1.0, qmin_ahi=-1.0, qmin_coord=0,
                lclose=False):
        prm = np.array([rmin, rmax, rmaxu, qmin, qmin_alo, qmin_ahi, dt], np.float64)
        counts = (C.c_int64 * 2)()
        n, m = self.pl_rh.shape[0], self.tp_rh.shape[0]
        lib().swo_discard(self.integ, n, m, self.GMcb, self.Rcb, self.cb_vb, self.pl_rh, self.pl_vh, self.pl_vb,
                          self.Gmass, self.radius, self.rhill, self.pl_status, self.nplenc, self.pl_isperi,
                          self.pl_atp, self.pl_peri, self.lfirst_peri, self.tp_rh, self.tp_vh, self.tp_vb,
                          self.tp_status, self.tp_isperi, self.tp_atp, self.tp_peri, prm, int(qmin_coord),
                          int(bool(lclose)), counts)
        return counts[0], counts[1]
