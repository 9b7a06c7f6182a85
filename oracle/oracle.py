"""ctypes binding of the test oracle (libvedo_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvedo_oracle.so")


class VoModel(C.Structure):
    _fields_ = [("shape", C.c_int), ("behavior", C.c_int), ("scope", C.c_int), ("group", C.c_int),
                ("density", C.c_double), ("density_factor", C.c_double),
                ("fext", C.c_double * 3), ("attr", C.c_double * 3), ("hole", C.c_double * 3)]


class VoIact(C.Structure):
    _fields_ = [("law", C.c_int), ("p", C.c_double * 8)]


class VoStatus(C.Structure):
    _fields_ = [("contact", C.c_ubyte), ("last_contact", C.c_ubyte), ("bond", C.c_ubyte),
                ("last_bond", C.c_ubyte), ("kn", C.c_double), ("init_vel", C.c_double),
                ("shear", C.c_double * 3), ("point", C.c_double * 3), ("dir", C.c_double * 3),
                ("to_master", C.c_double * 3), ("ang_to_master", C.c_double * 3),
                ("overlap", C.c_double)]


class VoScene(C.Structure):
    _fields_ = [("n_do", C.c_int), ("n_models", C.c_int), ("n_groups", C.c_int),
                ("models", C.POINTER(VoModel)), ("iact", C.POINTER(VoIact)),
                ("model_of", C.POINTER(C.c_int)),
                ("pos", C.POINTER(C.c_double)), ("vel", C.POINTER(C.c_double)),
                ("ox", C.POINTER(C.c_double)), ("oz", C.POINTER(C.c_double)),
                ("av", C.POINTER(C.c_double)),
                ("dt", C.c_double), ("t_start", C.c_double), ("t_stop", C.c_double),
                ("t_current", C.c_double), ("g", C.c_double * 3), ("safety_factor", C.c_double),
                ("up_iact", C.c_uint), ("record", C.c_uint), ("consultant", C.c_int),
                ("zoi_lo", C.c_double * 3), ("zoi_hi", C.c_double * 3),
                ("pbc_lo", C.c_double * 3), ("pbc_hi", C.c_double * 3)]


class VoContact(C.Structure):
    _fields_ = [("depth", C.c_double), ("n", C.c_double * 3), ("p", C.c_double * 3), ("active", C.c_int)]


STATUS_DTYPE = np.dtype([("contact", "u1"), ("last_contact", "u1"), ("bond", "u1"), ("last_bond", "u1"),
                         ("pad", "u1", 4), ("kn", "f8"), ("init_vel", "f8"), ("shear", "f8", 3),
                         ("point", "f8", 3), ("dir", "f8", 3), ("to_master", "f8", 3),
                         ("ang_to_master", "f8", 3), ("overlap", "f8")])
assert STATUS_DTYPE.itemsize == C.sizeof(VoStatus)

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        D = C.POINTER(C.c_double)
        U = C.POINTER(C.c_uint)
        L.vo_create.restype = C.c_void_p
        L.vo_create.argtypes = [C.POINTER(VoScene), C.c_char_p, C.c_int]
        L.vo_destroy.argtypes = [C.c_void_p]
        L.vo_step.argtypes = [C.c_void_p]
        L.vo_step_assigned.argtypes = [C.c_void_p, C.c_int, U, D, D, D, D, D]
        L.vo_set_history.argtypes = [C.c_void_p, C.c_int, U, U, C.POINTER(VoStatus)]
        L.vo_get_state.argtypes = [C.c_void_p, D, D, D, D, D, D, D]
        L.vo_time.restype = C.c_double
        L.vo_time.argtypes = [C.c_void_p]
        L.vo_system_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.vo_coordination_number.restype = C.c_double
        L.vo_coordination_number.argtypes = [C.c_void_p, D]
        L.vo_packing_density.restype = C.c_double
        L.vo_packing_density.argtypes = [C.c_void_p, D, C.c_int, C.c_int]
        L.vo_granular_temperature.argtypes = [C.c_void_p, C.c_int, D, D, D, D, D]
        for f in ("vo_num_interactions", "vo_num_candidates", "vo_num_active", "vo_num_do", "vo_num_records"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.vo_get_ids.argtypes = [C.c_void_p, U]
        L.vo_cleanup_stats.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        L.vo_get_interactions.argtypes = [C.c_void_p, U, U]
        L.vo_get_candidates.argtypes = [C.c_void_p, U, U]
        L.vo_get_active.argtypes = [C.c_void_p, U, U, C.POINTER(VoStatus)]
        L.vo_get_records.argtypes = [C.c_void_p, U, U]
        L.vo_cd_sphere_sphere.argtypes = [D, C.c_double, D, C.c_double, C.POINTER(VoContact)]
        L.vo_cd_sphere_plate.argtypes = [D, C.c_double, D, D, D, C.c_double, C.c_double, C.c_double,
                                         C.POINTER(VoContact)]
        L.vo_cd_sphere_cylinder.argtypes = [D, C.c_double, D, D, C.c_double, C.c_double, C.POINTER(VoContact)]
        L.vo_rotate_around.argtypes = [D, D, D]
        L.vo_solve.argtypes = [C.c_int, D, C.c_double, C.c_double, C.POINTER(VoContact),
                               D, D, D, D, D, D, C.c_double, C.POINTER(VoStatus), D, D, D]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint))


class OracleSim:
    """One CPU simulation of a WorldArrays scene (vedo_amd.world)."""

    def __init__(self, wa, consultant: int = 0):
        L = lib()
        self.n = wa.n_do
        M = len(wa.model_shape)
        G = wa.iact_law.shape[0]
        self._models = (VoModel * M)()
        for k in range(M):
            m = self._models[k]
            m.shape, m.behavior, m.scope, m.group = (int(wa.model_shape[k]), int(wa.model_behavior[k]),
                                                     int(wa.model_scope[k]), int(wa.model_group[k]))
            m.density, m.density_factor = float(wa.model_density[k]), float(wa.model_density_factor[k])
            hole = getattr(wa, "model_hole", None)
            for d in range(3):
                m.fext[d] = float(wa.model_fext[k, d])
                m.attr[d] = float(wa.model_attr[k, d])
                m.hole[d] = float(hole[k, d]) if hole is not None else 0.0
        self._iact = (VoIact * (G * G))()
        for i in range(G):
            for j in range(G):
                it = self._iact[i * G + j]
                it.law = int(wa.iact_law[i, j])
                for p in range(8):
                    it.p[p] = float(wa.iact_param[i, j, p])
        self._arrs = [np.ascontiguousarray(x, np.float64) for x in (wa.pos, wa.vel, wa.ox, wa.oz, wa.av)]
        self._model_of = np.ascontiguousarray(wa.model_of, np.int32)
        sc = VoScene()
        sc.n_do, sc.n_models, sc.n_groups = self.n, M, G
        sc.models, sc.iact = self._models, self._iact
        sc.model_of = self._model_of.ctypes.data_as(C.POINTER(C.c_int))
        sc.pos, sc.vel, sc.ox, sc.oz, sc.av = (_dp(a) for a in self._arrs)
        sc.dt, sc.t_start, sc.t_stop, sc.t_current = wa.dt, wa.t_start, wa.t_stop, wa.t_current
        for d in range(3):
            sc.g[d] = float(wa.g[d])
        sc.safety_factor = wa.safety_factor
        sc.up_iact, sc.record, sc.consultant = wa.up_iact, wa.record, consultant
        zlo, zhi = getattr(wa, "zoi_lo", np.zeros(3)), getattr(wa, "zoi_hi", np.zeros(3))
        for d in range(3):
            sc.zoi_lo[d], sc.zoi_hi[d] = float(zlo[d]), float(zhi[d])
        plo, phi = getattr(wa, "pbc_lo", np.zeros(3)), getattr(wa, "pbc_hi", np.zeros(3))
        for d in range(3):
            sc.pbc_lo[d], sc.pbc_hi[d] = float(plo[d]), float(phi[d])
        err = C.create_string_buffer(256)
        self.h = L.vo_create(C.byref(sc), err, 256)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.L = L

    def __del__(self):
        if getattr(self, "h", None):
            self.L.vo_destroy(self.h)
            self.h = None

    def step(self, k: int = 1):
        r = 1
        for _ in range(k):
            r = self.L.vo_step(self.h)
            if r < 0:
                raise RuntimeError("oracle step failed")
        return r

    def step_assigned(self, ids, pos, vel, ox, oz, av):
        ids = np.ascontiguousarray(ids, np.uint32)
        arrs = [np.ascontiguousarray(x, np.float64) for x in (pos, vel, ox, oz, av)]
        r = self.L.vo_step_assigned(self.h, len(ids), _up(ids), *(_dp(a) for a in arrs))
        if r < 0:
            raise RuntimeError("oracle step failed")
        return r

    @property
    def n_do(self) -> int:
        """Current number of DOs (the zone of interest erases DOs)."""
        return self.L.vo_num_do(self.h)

    def cleanup_stats(self):
        """Counts of Consultant::CleanUp so far: erased, moved, from_dead, replaced."""
        out = (C.c_longlong * 4)()
        self.L.vo_cleanup_stats(self.h, out)
        return dict(zip(("erased", "moved", "from_dead", "replaced"), list(out)))

    def ids(self):
        """The scene's original DO id of each current DO."""
        out = np.zeros(self.n_do, np.uint32)
        self.L.vo_get_ids(self.h, _up(out))
        return out

    def state(self):
        self.n = self.n_do
        out = {k: np.zeros((self.n, 3)) for k in ("pos", "vel", "ox", "oz", "av", "impact", "ang_impact")}
        self.L.vo_get_state(self.h, *(_dp(out[k]) for k in ("pos", "vel", "ox", "oz", "av", "impact", "ang_impact")))
        return out

    @property
    def time(self):
        return self.L.vo_time(self.h)

    def system_energy(self):
        """DOWorld::CalculateSystemEnergy of the current state, as a dict."""
        from vedo_amd.gpu import energy_dict
        out = np.zeros(16)
        self.L.vo_system_energy(self.h, _dp(out))
        return energy_dict(out)

    def coordination_number(self, box):
        """CalcCoordinateNum::computeAvgCoordinateNum over box (xmin xmax ymin ymax zmin zmax)."""
        return self.L.vo_coordination_number(self.h, _dp(np.ascontiguousarray(box, np.float64)))

    def packing_density(self, box, n_points: int = 20000, seed: int = 999):
        """CalcDensity::computeDensity over box (percent), njr generator seeded with `seed`."""
        return self.L.vo_packing_density(self.h, _dp(np.ascontiguousarray(box, np.float64)), int(n_points),
                                         int(seed))

    def granular_temperature(self, model: int, center, size, ox=(1.0, 0.0, 0.0), oz=(0.0, 0.0, 1.0)):
        """CalculateGranularTemperature of one model over a GSRectangle:
        (count, |mean v|, T_v, |mean w|, T_w)."""
        out = np.zeros(5)
        arr = [np.ascontiguousarray(x, np.float64) for x in (center, ox, oz, size)]
        self.L.vo_granular_temperature(self.h, int(model), _dp(arr[0]), _dp(arr[1]), _dp(arr[2]), _dp(arr[3]),
                                       _dp(out))
        return out

    def interactions(self):
        n = self.L.vo_num_interactions(self.h)
        s, m = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        self.L.vo_get_interactions(self.h, _up(s), _up(m))
        return s, m

    def candidates(self):
        n = self.L.vo_num_candidates(self.h)
        lo, hi = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        self.L.vo_get_candidates(self.h, _up(lo), _up(hi))
        return lo, hi

    def records(self):
        """Keys (lo, hi) of the IactRecordTab of the last rebuild or record step."""
        n = self.L.vo_num_records(self.h)
        lo, hi = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        if n:
            self.L.vo_get_records(self.h, _up(lo), _up(hi))
        return lo, hi

    def active(self):
        n = self.L.vo_num_active(self.h)
        lo, hi = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        st = np.zeros(n, STATUS_DTYPE)
        self.L.vo_get_active(self.h, _up(lo), _up(hi), st.ctypes.data_as(C.POINTER(VoStatus)))
        return lo, hi, st

    def set_history(self, lo, hi, st):
        lo = np.ascontiguousarray(lo, np.uint32)
        hi = np.ascontiguousarray(hi, np.uint32)
        st = np.ascontiguousarray(st, STATUS_DTYPE)
        r = self.L.vo_set_history(self.h, len(lo), _up(lo), _up(hi), st.ctypes.data_as(C.POINTER(VoStatus)))
        if r:
            raise RuntimeError("set_history failed")
