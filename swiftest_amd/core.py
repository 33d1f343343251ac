"""Python host mirror of the reference's operator API over the C ABI.

The functions keep the names and argument meaning of the reference's numeric
leaves (src/swiftest/swiftest_kick.f90, swiftest_drift.f90,
src/encounter/encounter_check.f90) and its in/out contract: accelerations are
accumulated into the caller's array, drift updates positions and velocities
in place and reports per-body flags, encounter checks return 1-based
(index1, index2, lvdotr) lists ordered index1 then index2.

Everything runs on the GPU through libswiftest_hip.so.  There is no CPU
fallback: if the library is missing or no GPU is visible, the first call
raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libswiftest_hip.so")

WHM, HELIO, SYMBA = 0, 1, 2
PAIR_EXACT, PAIR_FAST = 0, 1   # SWG_PAIR_* (pair factor of the direct sums)
# fields of an energy record (SWG_E_* in include/swiftest_hip.h)
ENERGY_FIELDS = ("ke_orbit", "ke_spin", "pe", "be", "be_cb", "te", "L_orbit_x", "L_orbit_y", "L_orbit_z",
                 "L_spin_x", "L_spin_y", "L_spin_z", "L_total_x", "L_total_y", "L_total_z", "GMtot")
ENERGY_NOUT = len(ENERGY_FIELDS)
ERRORS = {-1: "argument", -2: "hip", -3: "collision", -4: "drift", -5: "depth"}


class SwiftestHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"swiftest_hip error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class DiscardParams(C.Structure):
    _fields_ = [("rmin", C.c_double), ("rmax", C.c_double), ("rmaxu", C.c_double), ("qmin", C.c_double),
                ("qmin_alo", C.c_double), ("qmin_ahi", C.c_double), ("qmin_coord", C.c_int32),
                ("reserved", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("integrator", C.c_int32), ("lclose", C.c_int32),
                ("short_circuit", C.c_int32), ("istep_out", C.c_int32), ("rhill_present", C.c_int32),
                ("encounter_check", C.c_int32), ("pair_mode", C.c_int32), ("reserved", C.c_int32),
                ("gmtiny", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("nplpl_enc", C.c_int64), ("npltp_enc", C.c_int64), ("max_depth", C.c_int32),
                ("reserved", C.c_int32), ("nsubsteps", C.c_int64), ("ncollision", C.c_int64),
                ("ndrifterr", C.c_int64)]


_lib = None
_ctx = None

D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
vp = C.c_void_p


def library():
    """Load libswiftest_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise SwiftestHipError(-2, f"{LIBPATH} not built; run swiftest_amd/build.py")
        L = C.CDLL(LIBPATH)
        L.swg_last_error.restype = C.c_char_p
        L.swg_init.argtypes = [C.c_int, C.POINTER(vp)]
        L.swg_finalize.argtypes = [vp]
        L.swg_kick_getacch_int_all_tri_pl.argtypes = [vp, C.c_int32, C.c_int32, D, D, vp, D]
        L.swg_kick_getacch_int_all_flat_pl.argtypes = [vp, C.c_int32, C.c_int64, I32, D, D, vp, D]
        L.swg_kick_getacch_int_all_tp.argtypes = [vp, C.c_int32, C.c_int32, D, D, D, I32, D]
        L.swg_drift_all.argtypes = [vp, D, D, D, C.c_int32, C.c_double, I32, I32]
        L.swg_encounter_check_all_plpl.argtypes = [vp, C.c_int32, D, D, D, C.c_double, C.POINTER(C.c_int64)]
        L.swg_encounter_check_all_pltp.argtypes = [vp, C.c_int32, C.c_int32, D, D, D, D, D, C.c_double,
                                                   C.POINTER(C.c_int64)]
        L.swg_encounter_list_copy.argtypes = [vp, vp, vp, vp]
        L.swg_system_create.argtypes = [vp, C.POINTER(Params), C.c_int32, I32, I32, D, D, D, D, D, vp, D, D,
                                        C.POINTER(vp)]
        L.swg_system_run.argtypes = [vp, C.c_int64]
        L.swg_system_step.argtypes = [vp]
        L.swg_system_get.argtypes = [vp] * 10
        L.swg_system_get_misc.argtypes = [vp] * 6
        L.swg_system_stats.argtypes = [vp, C.POINTER(Stats)]
        L.swg_system_enc_list.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), vp, vp]
        L.swg_system_enc_status.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), vp]
        L.swg_system_time_kernels.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.swg_system_kernel_timing.argtypes = [vp, C.c_int32]
        L.swg_system_kernel_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                              C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.swg_system_destroy.argtypes = [vp]
        L.swg_system_accel.argtypes = [vp, vp, vp]
        L.swg_system_create_slots.argtypes = [vp, C.POINTER(Params), C.c_int32, I32, I32, C.POINTER(vp)]
        L.swg_system_load.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double] + [vp] * 7
        L.swg_system_get_slot.argtypes = [vp, C.c_int32] + [vp] * 9
        L.swg_system_slot_counts.argtypes = [vp, C.c_int32, vp, vp]
        L.swg_set_pair_mode.argtypes = [vp, C.c_int32]
        L.swg_system_set_pair_mode.argtypes = [vp, C.c_int32]
        L.swg_system_energy.argtypes = [vp, D, D, D, D]
        L.swg_system_discard.argtypes = [vp, C.POINTER(DiscardParams), D, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.swg_system_set_merge.argtypes = [vp, C.c_double, vp]
        L.swg_system_counts.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.swg_system_get_bodies.argtypes = [vp] * 5
        L.swg_system_pl_offsets.argtypes = [vp, vp]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise SwiftestHipError(rc, library().swg_last_error().decode())


def context(device: int = 0):
    """The process-wide library context on `device` (created on first use)."""
    global _ctx
    if _ctx is None:
        L = library()
        h = vp()
        _check(L.swg_init(device, C.byref(h)))
        _ctx = h
    return _ctx


def new_context(device: int = 0):
    """A further library context on `device`: its own HIP stream, so work
    issued through it overlaps with other contexts' work on the device."""
    L = library()
    h = vp()
    _check(L.swg_init(device, C.byref(h)))
    return h


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


# ---- leaf operators (same names and contracts as the reference) ----------

def set_pair_mode(mode):
    """Pair factor of the leaf operators: PAIR_EXACT (default) or PAIR_FAST."""
    _check(library().swg_set_pair_mode(context(), int(mode)))


def swiftest_kick_getacch_int_all_tri(npl, nplm, r, Gmass, radius, acc):
    """swiftest_kick_getacch_int_all_tri_rad_pl / _norad_pl (swiftest_kick.f90:165, :274).
    acc is accumulated into and returned."""
    acc = _c(acc).copy()
    rad = None if radius is None else _c(radius)
    _check(library().swg_kick_getacch_int_all_tri_pl(context(), int(npl), int(nplm), _c(r), _c(Gmass), _p(rad), acc))
    return acc


def swiftest_kick_getacch_int_all_flat(npl, k_plpl, r, Gmass, radius, acc):
    """swiftest_kick_getacch_int_all_flat_rad_pl / _norad_pl (swiftest_kick.f90:69, :118)."""
    acc = _c(acc).copy()
    k = _c(k_plpl, np.int32).reshape(-1, 2)
    rad = None if radius is None else _c(radius)
    _check(library().swg_kick_getacch_int_all_flat_pl(context(), int(npl), k.shape[0], k, _c(r), _c(Gmass),
                                                       _p(rad), acc))
    return acc


def swiftest_kick_getacch_int_all_tp(ntp, npl, rtp, rpl, GMpl, lmask, acc):
    """swiftest_kick_getacch_int_all_tp (swiftest_kick.f90:374)."""
    acc = _c(acc).copy()
    _check(library().swg_kick_getacch_int_all_tp(context(), int(ntp), int(npl), _c(rtp), _c(rpl), _c(GMpl),
                                                  _c(lmask, np.int32), acc))
    return acc


def swiftest_drift_all(mu, x, v, n, dt, lmask):
    """swiftest_drift_all (swiftest_drift.f90:60), GR off.  Returns (x, v, iflag)."""
    x = _c(x).copy()
    v = _c(v).copy()
    fl = np.zeros(int(n), np.int32)
    mu = np.broadcast_to(np.asarray(mu, np.float64), (int(n),)).copy()
    _check(library().swg_drift_all(context(), mu, x, v, int(n), float(dt), _c(lmask, np.int32), fl))
    return x, v, fl


def _take_list(n):
    a, b, c = (np.zeros(max(n, 1), np.int32) for _ in range(3))
    if n:
        _check(library().swg_encounter_list_copy(context(), _p(a), _p(b), _p(c)))
    return a[:n], b[:n], c[:n]


def encounter_check_all_plpl(npl, r, v, renc, dt):
    """encounter_check_all_plpl, TRIANGULAR (encounter_check.f90:14): (index1, index2, lvdotr)."""
    n = C.c_int64()
    _check(library().swg_encounter_check_all_plpl(context(), int(npl), _c(r), _c(v), _c(renc), float(dt),
                                                   C.byref(n)))
    return _take_list(n.value)


def encounter_check_all_pltp(npl, ntp, rpl, vpl, rtp, vtp, renc, dt):
    """encounter_check_all_pltp, TRIANGULAR (encounter_check.f90:112): (index1, index2, lvdotr)."""
    n = C.c_int64()
    _check(library().swg_encounter_check_all_pltp(context(), int(npl), int(ntp), _c(rpl), _c(vpl), _c(rtp),
                                                   _c(vtp), _c(renc), float(dt), C.byref(n)))
    return _take_list(n.value)


# ---- resident systems -----------------------------------------------------

class GpuSystem:
    """An ensemble of independent systems integrated on the GPU.

    `systems` is one swiftest_amd.ics.System or a list of them; all share the
    integrator and dt.  Body arrays of the systems are concatenated in order.
    """

    def __init__(self, integrator, systems, dt, lclose=True, short_circuit=True, istep_out=0, rhill=None, ctx=None,
                 encounter_check=0, pair_mode=PAIR_EXACT, gmtiny=0.0, merge=None):
        if not isinstance(systems, (list, tuple)):
            systems = [systems]
        self.L = library()
        self.nsys = len(systems)
        self.npl_s = np.array([s.npl for s in systems], np.int32)
        self.ntp_s = np.array([s.ntp for s in systems], np.int32)
        self.npl, self.ntp = int(self.npl_s.sum()), int(self.ntp_s.sum())
        cat = lambda key, w: _c(np.concatenate([getattr(s, key).reshape(-1, w) for s in systems])) \
            if systems else np.zeros((0, w))
        prm = Params(float(dt), int(integrator), int(lclose), int(short_circuit), int(istep_out),
                     int(rhill is not None), int(encounter_check), int(pair_mode), 0, float(gmtiny))
        rh = None if rhill is None else _c(rhill)
        h = vp()
        _check(self.L.swg_system_create(ctx if ctx is not None else context(), C.byref(prm), self.nsys,
                                        self.npl_s, self.ntp_s,
                                        _c([s.GMcb for s in systems]), cat("pl_rh", 3), cat("pl_vh", 3),
                                        _c(np.concatenate([s.Gmass for s in systems])),
                                        _c(np.concatenate([s.radius for s in systems])), _p(rh),
                                        cat("tp_rh", 3), cat("tp_vh", 3), C.byref(h)))
        self.h = h
        self._cb = (np.array([s.GMcb for s in systems], np.float64), np.array([s.Rcb for s in systems], np.float64))
        self._gm_in = [np.asarray(s.Gmass, np.float64) for s in systems]
        # merge = (GU, pl_mass or None): collisions resolved by merging inside
        # the step (swg_system_set_merge); the body count then shrinks
        self._merge_gu = None
        if merge is not None:
            GU, mass = merge
            self._mass_in = None if mass is None else _c(mass)
            _check(self.L.swg_system_set_merge(self.h, float(GU), _p(self._mass_in)))
            self._merge_gu = float(GU)

    @classmethod
    def slots(cls, integrator, npl_cap, ntp_cap, dt, lclose=True, short_circuit=True, ctx=None,
              encounter_check=0, pair_mode=PAIR_EXACT, gmtiny=0.0, merge_gu=None):
        """A slot pool (swg_system_create_slots): len(npl_cap) empty slots of
        the given capacities, all frozen until a system is loaded.  With
        merge_gu (G in the systems' units) collisions merge inside the step
        (swg_system_set_merge; masses Gmass / GU)."""
        self = cls.__new__(cls)
        self.L = library()
        self.nsys = len(npl_cap)
        self.npl_s = _c(npl_cap, np.int32)
        self.ntp_s = _c(ntp_cap, np.int32)
        self.npl, self.ntp = int(self.npl_s.sum()), int(self.ntp_s.sum())
        prm = Params(float(dt), int(integrator), int(lclose), int(short_circuit), 0, 0, int(encounter_check),
                     int(pair_mode), 0, float(gmtiny))
        h = vp()
        _check(self.L.swg_system_create_slots(ctx if ctx is not None else context(), C.byref(prm), self.nsys,
                                              self.npl_s, self.ntp_s, C.byref(h)))
        self.h = h
        self._cb = (np.ones(self.nsys), np.ones(self.nsys))
        self._gm_in = [np.zeros(n) for n in self.npl_s]
        self._merge_gu = None
        if merge_gu is not None:
            _check(self.L.swg_system_set_merge(self.h, float(merge_gu), None))
            self._merge_gu = float(merge_gu)
        return self

    def load(self, slot, system, rhill=None):
        """Put `system` (an ics.System) into `slot` (swg_system_load)."""
        rh = None if rhill is None else _c(rhill)
        _check(self.L.swg_system_load(self.h, int(slot), system.npl, system.ntp, float(system.GMcb),
                                      _p(_c(system.pl_rh).reshape(-1, 3)), _p(_c(system.pl_vh).reshape(-1, 3)),
                                      _p(_c(system.Gmass)), _p(_c(system.radius)), _p(rh),
                                      _p(_c(system.tp_rh).reshape(-1, 3)), _p(_c(system.tp_vh).reshape(-1, 3))))
        self._cb[0][slot], self._cb[1][slot] = system.GMcb, system.Rcb

    def slot_counts(self, slot):
        """Bodies the system in `slot` has now (swg_system_slot_counts)."""
        n, m = C.c_int32(), C.c_int32()
        _check(self.L.swg_system_slot_counts(self.h, int(slot), C.byref(n), C.byref(m)))
        return n.value, m.value

    def slot_bodies(self, slot):
        """Ids, masses, Gmass and radii of the massive bodies of the system in
        `slot`, in storage order."""
        b = self.bodies()
        off = np.zeros(self.nsys + 1, np.int32)
        _check(self.L.swg_system_pl_offsets(self.h, _p(off)))
        n, _ = self.slot_counts(slot)
        return {k: v[off[slot]:off[slot] + n] for k, v in b.items()}

    def get_slot(self, slot):
        """State of the system `slot` holds (swg_system_get_slot)."""
        n, m = self.slot_counts(slot)
        out = dict(pl_rh=np.zeros((n, 3)), pl_vh=np.zeros((n, 3)), pl_vb=np.zeros((n, 3)), tp_rh=np.zeros((m, 3)),
                   tp_vh=np.zeros((m, 3)), tp_vb=np.zeros((m, 3)), pl_status=np.zeros(n, np.int32),
                   tp_status=np.zeros(m, np.int32), pl_id=np.zeros(n, np.int32))
        _check(self.L.swg_system_get_slot(self.h, int(slot), *[_p(out[k]) for k in (
            "pl_rh", "pl_vh", "pl_vb", "tp_rh", "tp_vh", "tp_vb", "pl_status", "tp_status", "pl_id")]))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self.L.swg_system_destroy(self.h)
            self.h = None

    def run(self, nsteps):
        _check(self.L.swg_system_run(self.h, int(nsteps)))

    def step(self):
        _check(self.L.swg_system_step(self.h))

    def _sync_counts(self):
        n, m, k = C.c_int32(), C.c_int32(), C.c_int64()
        _check(self.L.swg_system_counts(self.h, C.byref(n), C.byref(m), C.byref(k)))
        off = np.zeros(self.nsys + 1, np.int32)
        _check(self.L.swg_system_pl_offsets(self.h, _p(off)))
        self.npl_s = np.diff(off).astype(np.int32)
        self.npl, self.ntp = n.value, m.value
        return k.value

    @property
    def nmerge(self):
        """Merges resolved so far (swg_system_counts)."""
        return self._sync_counts()

    def bodies(self):
        """Ids, masses, Gmass and radii of the massive bodies, in storage order."""
        self._sync_counts()
        n = self.npl
        out = dict(id=np.zeros(n, np.int32), mass=np.zeros(n), Gmass=np.zeros(n), radius=np.zeros(n))
        _check(self.L.swg_system_get_bodies(self.h, *[_p(out[k]) for k in ("id", "mass", "Gmass", "radius")]))
        return out

    def state(self):
        self._sync_counts()
        npl, ntp = self.npl, self.ntp
        out = dict(pl_rh=np.zeros((npl, 3)), pl_vh=np.zeros((npl, 3)), pl_vb=np.zeros((npl, 3)),
                   tp_rh=np.zeros((ntp, 3)), tp_vh=np.zeros((ntp, 3)), tp_vb=np.zeros((ntp, 3)),
                   pl_status=np.zeros(npl, np.int32), tp_status=np.zeros(ntp, np.int32),
                   pl_id=np.zeros(npl, np.int32))
        _check(self.L.swg_system_get(self.h, *[_p(out[k]) for k in (
            "pl_rh", "pl_vh", "pl_vb", "tp_rh", "tp_vh", "tp_vb", "pl_status", "tp_status", "pl_id")]))
        rhill = np.zeros(npl)
        cb_vb, ptb, pte = (np.zeros((self.nsys, 3)) for _ in range(3))
        lf = np.zeros(self.nsys, np.int32)
        _check(self.L.swg_system_get_misc(self.h, _p(rhill), _p(cb_vb), _p(ptb), _p(pte), _p(lf)))
        out.update(rhill=rhill, cb_vb=cb_vb, ptbeg=ptb, ptend=pte, lfirst=lf)
        return out

    def stats(self):
        st = Stats()
        _check(self.L.swg_system_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in Stats._fields_ if k != "reserved"}

    def enc_list(self, which=0):
        n = C.c_int64()
        _check(self.L.swg_system_enc_list(self.h, which, C.byref(n), None, None))
        a, b = np.zeros(max(n.value, 1), np.int32), np.zeros(max(n.value, 1), np.int32)
        _check(self.L.swg_system_enc_list(self.h, which, C.byref(n), _p(a), _p(b)))
        return a[:n.value], b[:n.value]

    def enc_status(self, which=0):
        """Status of each pair of the level-0 list of the last step (0 ACTIVE,
        -14 COLLIDED), in list order (swg_system_enc_status)."""
        n = C.c_int64()
        _check(self.L.swg_system_enc_status(self.h, which, C.byref(n), None))
        st = np.zeros(max(n.value, 1), np.int32)
        _check(self.L.swg_system_enc_status(self.h, which, C.byref(n), _p(st)))
        return st[:n.value]

    def accel(self):
        """(pl_acc, tp_acc): the direct-sum accelerations at the current
        positions from the step's own kernels (swg_system_accel)."""
        self._sync_counts()
        a, b = np.zeros((self.npl, 3)), np.zeros((self.ntp, 3))
        _check(self.L.swg_system_accel(self.h, _p(a), _p(b)))
        return a, b

    def set_pair_mode(self, mode):
        """PAIR_EXACT or PAIR_FAST for the following steps."""
        _check(self.L.swg_system_set_pair_mode(self.h, int(mode)))

    def kernel_timing(self, enable=True):
        """Record HIP events around every acceleration-kernel launch of the
        following steps (enabling resets the record)."""
        _check(self.L.swg_system_kernel_timing(self.h, int(enable)))

    def kernel_times(self):
        """(avg ms of the pl-tp kernel, launches, avg ms of the pl-pl kernel, launches)."""
        a, b = C.c_double(), C.c_double()
        na, nb = C.c_int64(), C.c_int64()
        _check(self.L.swg_system_kernel_times(self.h, C.byref(a), C.byref(na), C.byref(b), C.byref(nb)))
        return a.value, na.value, b.value, nb.value

    def use_graphs(self, enable=True):
        """Replay the step's fixed launch sequences from captured HIP graphs."""
        _check(self.L.swg_system_use_graphs(self.h, int(enable)))

    def time_encounters(self, reps=3):
        """(ms per level-0 pl-pl check, pairs found, broad-phase candidates or -1)."""
        ms, npairs, ncand = C.c_double(), C.c_int64(), C.c_int64()
        _check(self.L.swg_system_time_encounters(self.h, int(reps), C.byref(ms), C.byref(npairs), C.byref(ncand)))
        return ms.value, npairs.value, ncand.value

    def energy(self, cb_mass=None, cb_radius=None, pl_mass=None, GU=None):
        """Energy and angular momentum of every system (swg_system_energy):
        an (nsys, ENERGY_NOUT) array with the fields named in ENERGY_FIELDS.
        Masses default to Gmass / GU (GU: G in the system's units, by default
        that of swiftest_amd.ics), pl_mass given in the stored body order."""
        from .ics import GU as GU_DEFAULT
        gu = GU_DEFAULT if GU is None else GU
        if cb_mass is None:
            cb_mass = self._cb[0] / gu
        if cb_radius is None:
            cb_radius = self._cb[1]
        self._sync_counts()
        if pl_mass is None and getattr(self, "_merge_gu", None) is not None:
            pl_mass = self.bodies()["mass"]
        if pl_mass is None:
            ids = np.zeros(self.npl, np.int32)
            _check(self.L.swg_system_get(self.h, *([None] * 8), _p(ids)))
            pl_mass = np.zeros(self.npl)
            o = 0
            for s, n in enumerate(self.npl_s):
                pl_mass[o:o + n] = self._gm_in[s][ids[o:o + n] - 1] / gu
                o += n
        out = np.zeros((self.nsys, ENERGY_NOUT))
        _check(self.L.swg_system_energy(self.h, _c(cb_mass), _c(cb_radius), _c(pl_mass), out))
        return out

    def discard(self, rmin=-1.0, rmax=-1.0, rmaxu=-1.0, qmin=-1.0, qmin_alo=-1.0, qmin_ahi=-1.0, qmin_coord=0,
                cb_radius=None):
        """The reference's discard checks after a step (swg_system_discard);
        returns (massive bodies, test particles) newly discarded.  qmin_coord:
        0 HELIO, 1 BARY.  Discarded bodies get their status in place."""
        dp = DiscardParams(float(rmin), float(rmax), float(rmaxu), float(qmin), float(qmin_alo), float(qmin_ahi),
                           int(qmin_coord), 0)
        rcb = self._cb[1] if cb_radius is None else cb_radius
        a, b = C.c_int64(), C.c_int64()
        _check(self.L.swg_system_discard(self.h, C.byref(dp), _c(rcb), C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_frozen(self, frozen):
        """frozen[s] = 1 leaves system s out of the following steps."""
        f = np.ascontiguousarray(np.asarray(frozen, np.int32))
        assert f.shape == (self.nsys,)
        _check(self.L.swg_system_set_frozen(self.h, _p(f)))

    def autotune(self, reps=3):
        """Pick the pair kernels' workgroup sizes by timing; returns them."""
        b_pl, b_tp = C.c_int32(), C.c_int32()
        _check(self.L.swg_system_autotune(self.h, int(reps), C.byref(b_pl), C.byref(b_tp)))
        return b_pl.value, b_tp.value

    def time_kernels(self, reps=5):
        tp_ms, pl_ms = C.c_double(), C.c_double()
        _check(self.L.swg_system_time_kernels(self.h, int(reps), C.byref(tp_ms), C.byref(pl_ms)))
        return tp_ms.value, pl_ms.value
