"""ctypes binding of libvedo_hip.so (include/vedo_gpu.h).

``GpuSimulation`` is the GPU counterpart of running ``anne`` on a world:
create (Consultant::Reset + SimMediator::Initiate), step
(SimMediator::Run), read back DOStatus and the IactRecordTab payload.
There is no CPU fallback: if the HIP library is missing or no GPU is
present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .world import QUASIPLATE_HOLE

HERE = os.path.dirname(os.path.abspath(__file__))
# VG_LIBRARY: another build of the same library (A/B timing of build variants)
LIB_PATH = os.environ.get("VG_LIBRARY") or os.path.join(HERE, "libvedo_hip.so")


class VgModel(C.Structure):
    _fields_ = [("shape", C.c_int), ("behavior", C.c_int), ("scope", C.c_int), ("group", C.c_int),
                ("density", C.c_double), ("density_factor", C.c_double),
                ("fext", C.c_double * 3), ("attr", C.c_double * 6)]


class VgIact(C.Structure):
    _fields_ = [("law", C.c_int), ("p", C.c_double * 8)]


class VgStatus(C.Structure):
    _fields_ = [("contact", C.c_ubyte), ("last_contact", C.c_ubyte), ("bond", C.c_ubyte),
                ("last_bond", C.c_ubyte), ("kn", C.c_double), ("init_vel", C.c_double),
                ("shear", C.c_double * 3), ("point", C.c_double * 3), ("dir", C.c_double * 3),
                ("to_master", C.c_double * 3), ("ang_to_master", C.c_double * 3),
                ("overlap", C.c_double)]


_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint)


class VgSceneDesc(C.Structure):
    _fields_ = [("n_do", C.c_int), ("n_models", C.c_int), ("n_groups", C.c_int),
                ("models", C.POINTER(VgModel)), ("iact", C.POINTER(VgIact)),
                ("model_of", C.POINTER(C.c_int)),
                ("pos", _D), ("vel", _D), ("ox", _D), ("oz", _D), ("av", _D),
                ("dt", C.c_double), ("t_start", C.c_double), ("t_stop", C.c_double),
                ("t_current", C.c_double), ("g", C.c_double * 3), ("safety_factor", C.c_double),
                ("up_iact", C.c_uint), ("record", C.c_uint), ("consultant", C.c_int),
                ("inner_skin", C.c_double), ("zoi_lo", C.c_double * 3), ("zoi_hi", C.c_double * 3),
                ("pbc_lo", C.c_double * 3), ("pbc_hi", C.c_double * 3)]


class VgBatchDesc(C.Structure):
    _fields_ = [("max_spheres", C.c_longlong), ("max_walls", C.c_int), ("max_scenes", C.c_int),
                ("neighbour_capacity", C.c_int), ("up_iact", C.c_uint)]


STATUS_DTYPE = np.dtype([("contact", "u1"), ("last_contact", "u1"), ("bond", "u1"), ("last_bond", "u1"),
                         ("pad", "u1", 4), ("kn", "f8"), ("init_vel", "f8"), ("shear", "f8", 3),
                         ("point", "f8", 3), ("dir", "f8", 3), ("to_master", "f8", 3),
                         ("ang_to_master", "f8", 3), ("overlap", "f8")])
assert STATUS_DTYPE.itemsize == C.sizeof(VgStatus)

EXPORTS = ("vg_create", "vg_destroy", "vg_last_error", "vg_create_error", "vg_set_history", "vg_step",
           "vg_sync", "vg_step_assigned", "vg_set_wall_rotation", "vg_get_state", "vg_time",
           "vg_num_interactions", "vg_scene_num_do", "vg_scene_models", "vg_set_skin", "vg_set_skin_auto", "vg_get_skin", "vg_set_precision", "vg_get_precision", "vg_get_active", "vg_kernel_timing", "vg_step_timing", "vg_rebuild_timing", "vg_check_lists", "vg_stats", "vg_copy_bench",
           "vg_system_energy", "vg_batch_create", "vg_admit", "vg_retire", "vg_scene_status",
           "vg_scene_get_state", "vg_scene_get_active", "vg_scene_energy", "vg_scene_set_wall_rotation",
           "vg_batch_info", "vg_set_graphs", "vg_graph_replays", "vg_scene_coordination_number", "vg_scene_packing_density",
           "vg_scene_packing_density_noalloc", "vg_scene_granular_temperature")

_lib = None

# Layout of vg_system_energy's output (DOWorld::CalculateSystemEnergy,
# DOWorld.cpp:481-632, the values it stores in SystemParameter).
PRECISION_FP64, PRECISION_MIXED = 0, 1   # vg_set_precision

ENERGY_FIELDS = ("potential", "translation", "rotation", "momentum_avg", "angular_momentum_avg", "momentum_norm",
                 "angular_momentum_norm", "velocity_max", "velocity_min", "angular_velocity_max",
                 "angular_velocity_min", "mobile")


def energy_dict(out):
    o = np.asarray(out, np.float64)
    return {"potential": o[0], "translation": o[1], "rotation": o[2], "momentum_avg": o[3:6].copy(),
            "angular_momentum_avg": o[6:9].copy(), "momentum_norm": o[9], "angular_momentum_norm": o[10],
            "velocity_max": o[11], "velocity_min": o[12], "angular_velocity_max": o[13],
            "angular_velocity_min": o[14], "mobile": int(o[15])}


class GpuError(RuntimeError):
    pass


def lib():
    """Load libvedo_hip.so; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpuError(f"{LIB_PATH} is not built: run `python -m vedo_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.vg_create.argtypes = [C.POINTER(VgSceneDesc), C.POINTER(vp)]
        L.vg_destroy.argtypes = [vp]
        L.vg_last_error.restype = C.c_char_p
        L.vg_last_error.argtypes = [vp]
        L.vg_create_error.restype = C.c_char_p
        L.vg_set_history.argtypes = [vp, C.c_int, _U, _U, C.POINTER(VgStatus)]
        L.vg_step.argtypes = [vp, C.c_int]
        L.vg_sync.argtypes = [vp]
        L.vg_step_assigned.argtypes = [vp, C.c_int, _U, _D, _D, _D, _D, _D]
        L.vg_set_wall_rotation.argtypes = [vp, C.c_int, _U, _D, _D, C.c_double, C.c_double]
        L.vg_get_state.argtypes = [vp, _D, _D, _D, _D, _D, _D, _D]
        L.vg_time.restype = C.c_double
        L.vg_time.argtypes = [vp]
        L.vg_num_interactions.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.vg_scene_num_do.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        L.vg_scene_models.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        L.vg_get_active.argtypes = [vp, C.POINTER(C.c_longlong), _U, _U, C.POINTER(VgStatus)]
        L.vg_step_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.vg_kernel_timing.argtypes = [vp, C.c_int]
        L.vg_rebuild_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.vg_check_lists.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.vg_set_skin.argtypes = [vp, C.c_double]
        L.vg_scene_coordination_number.argtypes = [vp, C.c_int, _D, _D]
        L.vg_scene_packing_density.argtypes = [vp, C.c_int, _D, C.c_int, C.c_int, _D]
        L.vg_scene_packing_density_noalloc.argtypes = [vp, C.c_int, _D, C.c_int, C.c_int, _D]
        L.vg_scene_granular_temperature.argtypes = [vp, C.c_int, C.c_int, _D, _D, _D, _D, _D]
        L.vg_set_skin_auto.argtypes = [vp, C.c_double, C.c_double]
        L.vg_get_skin.restype = C.c_double
        L.vg_get_skin.argtypes = [vp]
        L.vg_set_precision.argtypes = [vp, C.c_int]
        L.vg_get_precision.argtypes = [vp]
        L.vg_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_int),
                               C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
        L.vg_copy_bench.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.vg_system_energy.argtypes = [vp, _D]
        L.vg_batch_create.argtypes = [C.POINTER(VgBatchDesc), C.POINTER(vp)]
        L.vg_admit.argtypes = [vp, C.POINTER(VgSceneDesc), C.c_longlong, C.POINTER(C.c_int)]
        L.vg_retire.argtypes = [vp, C.c_int]
        L.vg_scene_status.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong), _D]
        L.vg_scene_get_state.argtypes = [vp, C.c_int, _D, _D, _D, _D, _D]
        L.vg_scene_get_active.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong), _U, _U, C.POINTER(VgStatus)]
        L.vg_scene_energy.argtypes = [vp, C.c_int, _D]
        L.vg_scene_set_wall_rotation.argtypes = [vp, C.c_int, C.c_int, _U, _D, _D, C.c_double, C.c_double]
        L.vg_set_graphs.argtypes = [vp, C.c_int]
        L.vg_graph_replays.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.vg_batch_info.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_ulonglong)]
        _lib = L
    return _lib


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 10) -> float:
    """Achieved read+write GB/s of a plain streaming copy kernel."""
    gbps = C.c_double()
    rc = lib().vg_copy_bench(nbytes, iters, C.byref(gbps))
    if rc:
        raise GpuError(f"vg_copy_bench failed ({rc})")
    return gbps.value


def _dp(a):
    return a.ctypes.data_as(_D)


def scene_desc(wa, inner_skin: float = 0.0, consultant: int = 0):
    """vg_scene_desc of a WorldArrays; returns (desc, buffers to keep alive).
    consultant: 0 LeapConsultant's candidate list, 1 NBSConsultant's."""
    M = len(wa.model_shape)
    G = wa.iact_law.shape[0]
    models = (VgModel * M)()
    for k in range(M):
        m = models[k]
        m.shape, m.behavior, m.scope, m.group = (int(wa.model_shape[k]), int(wa.model_behavior[k]),
                                                 int(wa.model_scope[k]), int(wa.model_group[k]))
        m.density, m.density_factor = float(wa.model_density[k]), float(wa.model_density_factor[k])
        for d in range(3):
            m.fext[d] = float(wa.model_fext[k, d])
            m.attr[d] = float(wa.model_attr[k, d])
            # QuasiPlateWithCircularHole: hole radius, x / y offset
            m.attr[3 + d] = float(wa.model_hole[k, d]) if int(wa.model_shape[k]) == QUASIPLATE_HOLE else 0.0
    iact = (VgIact * (G * G))()
    for i in range(G):
        for j in range(G):
            it = iact[i * G + j]
            it.law = int(wa.iact_law[i, j])
            for p in range(8):
                it.p[p] = float(wa.iact_param[i, j, p])
    arrs = [np.ascontiguousarray(x, np.float64) for x in (wa.pos, wa.vel, wa.ox, wa.oz, wa.av)]
    model_of = np.ascontiguousarray(wa.model_of, np.int32)
    d = VgSceneDesc()
    d.n_do, d.n_models, d.n_groups = wa.n_do, M, G
    d.models, d.iact = models, iact
    d.model_of = model_of.ctypes.data_as(C.POINTER(C.c_int))
    d.pos, d.vel, d.ox, d.oz, d.av = (_dp(a) for a in arrs)
    d.dt, d.t_start, d.t_stop, d.t_current = wa.dt, wa.t_start, wa.t_stop, wa.t_current
    for c in range(3):
        d.g[c] = float(wa.g[c])
    d.safety_factor = wa.safety_factor
    d.up_iact, d.record, d.consultant = wa.up_iact, wa.record, int(consultant)
    d.inner_skin = inner_skin
    zlo, zhi = getattr(wa, "zoi_lo", np.zeros(3)), getattr(wa, "zoi_hi", np.zeros(3))
    for k in range(3):
        d.zoi_lo[k], d.zoi_hi[k] = float(zlo[k]), float(zhi[k])
    plo, phi = getattr(wa, "pbc_lo", np.zeros(3)), getattr(wa, "pbc_hi", np.zeros(3))
    for k in range(3):
        d.pbc_lo[k], d.pbc_hi[k] = float(plo[k]), float(phi[k])
    return d, (models, iact, arrs, model_of)


def _up(a):
    return a.ctypes.data_as(_U)


class _Knight:
    """Knight analyses (src/vedo/knight) of a scene's current state, on the device."""

    def coordination_number(self, box, scene: int = 0) -> float:
        """CalcCoordinateNum::computeAvgCoordinateNum over box (xmin xmax ymin ymax zmin zmax)."""
        b = np.ascontiguousarray(box, np.float64)
        out = C.c_double()
        self._check(self.L.vg_scene_coordination_number(self.h, int(scene), _dp(b), C.byref(out)))
        return out.value

    def packing_density(self, box, n_points: int = 20000, seed: int = 999, scene: int = 0) -> float:
        """CalcDensity::computeDensity over box, in percent."""
        b = np.ascontiguousarray(box, np.float64)
        out = C.c_double()
        self._check(self.L.vg_scene_packing_density(self.h, int(scene), _dp(b), int(n_points), int(seed),
                                                    C.byref(out)))
        return out.value

    def granular_temperature(self, model: int, center, size, ox=(1.0, 0.0, 0.0), oz=(0.0, 0.0, 1.0),
                             scene: int = 0):
        """CalculateGranularTemperature of one sphere model over a GSRectangle:
        (count, |mean v|, T_v, |mean w|, T_w)."""
        arr = [np.ascontiguousarray(x, np.float64) for x in (center, ox, oz, size)]
        out = np.zeros(5)
        self._check(self.L.vg_scene_granular_temperature(self.h, int(scene), int(model), _dp(arr[0]), _dp(arr[1]),
                                                         _dp(arr[2]), _dp(arr[3]), _dp(out)))
        return out


class GpuSimulation(_Knight):
    """One scene on the current GPU (a WorldArrays from vedo_amd.world)."""

    def __init__(self, wa, inner_skin: float = 0.0, consultant: int = 0):
        L = lib()
        self.L = L
        self.n = wa.n_do
        d, self._keep = scene_desc(wa, inner_skin, consultant)
        h = C.c_void_p()
        rc = L.vg_create(C.byref(d), C.byref(h))
        if rc:
            raise GpuError(f"vg_create failed ({rc}): {L.vg_create_error().decode()}")
        self.h = h

    def _check(self, rc):
        if rc:
            raise GpuError(f"libvedo_hip error {rc}: {self.L.vg_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.vg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, k: int = 1, sync: bool = True):
        self._check(self.L.vg_step(self.h, int(k)))
        if sync:
            self.sync()

    def sync(self):
        self._check(self.L.vg_sync(self.h))

    def step_assigned(self, ids, pos, vel, ox, oz, av):
        ids = np.ascontiguousarray(ids, np.uint32)
        arrs = [np.ascontiguousarray(x, np.float64) for x in (pos, vel, ox, oz, av)]
        self._check(self.L.vg_step_assigned(self.h, len(ids), _up(ids), *(_dp(a) for a in arrs)))

    def system_energy(self):
        """DOWorld::CalculateSystemEnergy of the current state (ENERGY_FIELDS)."""
        out = np.zeros(16)
        self._check(self.L.vg_system_energy(self.h, _dp(out)))
        return energy_dict(out)

    def set_wall_rotation(self, ids, axis_point, axis_dir, omega: float, t_ref: float = 0.0):
        ids = np.ascontiguousarray(ids, np.uint32)
        c = np.ascontiguousarray(axis_point, np.float64)
        a = np.ascontiguousarray(axis_dir, np.float64)
        self._check(self.L.vg_set_wall_rotation(self.h, len(ids), _up(ids), _dp(c), _dp(a), omega, t_ref))

    @property
    def n_do(self) -> int:
        """Current number of DOs (the zone of interest erases DOs)."""
        n = C.c_int()
        self._check(self.L.vg_scene_num_do(self.h, 0, C.byref(n)))
        return n.value

    def models(self):
        """DOModel index of each remaining DO (the DOName of its DOStatus)."""
        out = np.zeros(self.n_do, np.int32)
        self._check(self.L.vg_scene_models(self.h, 0, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def state(self):
        self.n = self.n_do
        out = {k: np.zeros((self.n, 3)) for k in ("pos", "vel", "ox", "oz", "av", "impact", "ang_impact")}
        self._check(self.L.vg_get_state(self.h, *(_dp(out[k]) for k in
                                                   ("pos", "vel", "ox", "oz", "av", "impact", "ang_impact"))))
        return out

    @property
    def time(self):
        return self.L.vg_time(self.h)

    def num_interactions(self):
        n = C.c_longlong()
        self._check(self.L.vg_num_interactions(self.h, C.byref(n)))
        return n.value

    def count_active(self) -> int:
        n = C.c_longlong()
        self._check(self.L.vg_get_active(self.h, C.byref(n), None, None, None))
        return n.value

    def active(self):
        n = C.c_longlong(self.count_active())
        lo, hi = np.zeros(n.value, np.uint32), np.zeros(n.value, np.uint32)
        st = np.zeros(n.value, STATUS_DTYPE)
        self._check(self.L.vg_get_active(self.h, C.byref(n), _up(lo), _up(hi),
                                         st.ctypes.data_as(C.POINTER(VgStatus))))
        return lo[:n.value], hi[:n.value], st[:n.value]

    def set_history(self, lo, hi, st):
        """Restore IactRecordTab statuses before the first step (vg_set_history)."""
        lo = np.ascontiguousarray(lo, np.uint32)
        hi = np.ascontiguousarray(hi, np.uint32)
        st = np.ascontiguousarray(st, STATUS_DTYPE)
        self._check(self.L.vg_set_history(self.h, len(lo), _up(lo), _up(hi),
                                          st.ctypes.data_as(C.POINTER(VgStatus))))

    def set_skin(self, skin: float):
        self._check(self.L.vg_set_skin(self.h, float(skin)))

    def set_skin_auto(self, floor: float, cap: float):
        """Skin per scene from its speed at every rebuild (vg_set_skin_auto)."""
        self._check(self.L.vg_set_skin_auto(self.h, float(floor), float(cap)))

    @property
    def skin(self) -> float:
        return self.L.vg_get_skin(self.h)

    def set_precision(self, mode: int):
        """PRECISION_FP64 (the reference's arithmetic) or PRECISION_MIXED."""
        self._check(self.L.vg_set_precision(self.h, int(mode)))

    @property
    def precision(self) -> int:
        return self.L.vg_get_precision(self.h)

    def kernel_timing(self, max_launches: int):
        self._check(self.L.vg_kernel_timing(self.h, int(max_launches)))

    def timing(self):
        ms, k = C.c_double(), C.c_int()
        self._check(self.L.vg_step_timing(self.h, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    def rebuild_timing(self):
        """(summed ms, count) of the rebuilds since kernel_timing()."""
        ms, k = C.c_double(), C.c_int()
        self._check(self.L.vg_rebuild_timing(self.h, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    def stats(self):
        rb, rl, mf, cap, cs = C.c_longlong(), C.c_longlong(), C.c_int(), C.c_int(), C.c_longlong()
        self._check(self.L.vg_stats(self.h, C.byref(rb), C.byref(rl), C.byref(mf), C.byref(cap), C.byref(cs)))
        return dict(rebuilds=rb.value, relists=rl.value, max_fill=mf.value, capacity=cap.value,
                    contact_steps=cs.value)

    def check_lists(self):
        """vg_check_lists: violations by kind and the first one."""
        return _check_lists(self)


def _check_lists(obj):
    out = (C.c_longlong * 11)()
    obj._check(obj.L.vg_check_lists(obj.h, out))
    v = list(out)
    names = ("model", "scene", "neighbour_count", "neighbour_entry", "history_slot", "candidate_count",
             "candidate_entry", "constrained_slot")
    return {"violations": {names[k]: v[k] for k in range(8)}, "first": v[8:11]}


SCENE_FREE, SCENE_QUEUED, SCENE_RUNNING, SCENE_FINISHED, SCENE_RETIRING, SCENE_DONE, SCENE_FAILED = range(7)


class GpuBatch(_Knight):
    """Many independent scenes on the current GPU, stepped together (one
    kernel launch per step for all of them).  Scenes are admitted with a step
    count, placed on the device at the next rebuild boundary, frozen when
    they reach their step count and moved to the host at the boundary after;
    each scene's result equals running it alone."""

    def __init__(self, max_spheres: int, max_walls: int = 4096, max_scenes: int = 1024, up_iact: int = 20,
                 neighbour_capacity: int = 0):
        L = lib()
        self.L = L
        bd = VgBatchDesc()
        bd.max_spheres, bd.max_walls, bd.max_scenes = int(max_spheres), int(max_walls), int(max_scenes)
        bd.neighbour_capacity, bd.up_iact = int(neighbour_capacity), int(up_iact)
        h = C.c_void_p()
        rc = L.vg_batch_create(C.byref(bd), C.byref(h))
        if rc:
            raise GpuError(f"vg_batch_create failed ({rc}): {L.vg_create_error().decode()}")
        self.h = h
        self.n_do = {}

    def _check(self, rc):
        if rc:
            raise GpuError(f"libvedo_hip error {rc}: {self.L.vg_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.vg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def admit(self, wa, steps: int = -1, inner_skin: float = 0.0) -> int:
        d, keep = scene_desc(wa, inner_skin)
        sid = C.c_int()
        self._check(self.L.vg_admit(self.h, C.byref(d), int(steps), C.byref(sid)))
        del keep
        self.n_do[sid.value] = wa.n_do
        return sid.value

    def retire(self, scene: int):
        self._check(self.L.vg_retire(self.h, int(scene)))

    def step(self, k: int = 1, sync: bool = True):
        self._check(self.L.vg_step(self.h, int(k)))
        if sync:
            self.sync()

    def sync(self):
        self._check(self.L.vg_sync(self.h))

    def status(self, scene: int):
        st, n, t = C.c_int(), C.c_longlong(), C.c_double()
        self._check(self.L.vg_scene_status(self.h, int(scene), C.byref(st), C.byref(n), C.byref(t)))
        return st.value, n.value, t.value

    def models(self, scene: int):
        """DOModel index of each remaining DO of a scene (the DOName of its DOStatus)."""
        m = C.c_int()
        self._check(self.L.vg_scene_num_do(self.h, int(scene), C.byref(m)))
        out = np.zeros(m.value, np.int32)
        self._check(self.L.vg_scene_models(self.h, int(scene), out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def state(self, scene: int):
        m = C.c_int()
        self._check(self.L.vg_scene_num_do(self.h, int(scene), C.byref(m)))
        n = self.n_do[scene] = m.value
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "ox", "oz", "av")}
        self._check(self.L.vg_scene_get_state(self.h, int(scene), *(_dp(out[k]) for k in
                                                                    ("pos", "vel", "ox", "oz", "av"))))
        return out

    def active(self, scene: int):
        n = C.c_longlong(0)
        self._check(self.L.vg_scene_get_active(self.h, int(scene), C.byref(n), None, None, None))
        cap = n.value
        lo = np.zeros(cap, np.uint32)
        hi = np.zeros(cap, np.uint32)
        st = np.zeros(cap, STATUS_DTYPE)
        n = C.c_longlong(cap)
        self._check(self.L.vg_scene_get_active(self.h, int(scene), C.byref(n), _up(lo), _up(hi),
                                               st.ctypes.data_as(C.POINTER(VgStatus))))
        return lo[:n.value], hi[:n.value], st[:n.value]

    def energy(self, scene: int):
        out = np.zeros(16)
        self._check(self.L.vg_scene_energy(self.h, int(scene), _dp(out)))
        return energy_dict(out)

    def set_wall_rotation(self, scene: int, ids, axis_point, axis_dir, omega: float, t_ref: float = 0.0):
        ids = np.ascontiguousarray(ids, np.uint32)
        c = np.ascontiguousarray(axis_point, np.float64)
        a = np.ascontiguousarray(axis_dir, np.float64)
        self._check(self.L.vg_scene_set_wall_rotation(self.h, int(scene), len(ids), _up(ids), _dp(c), _dp(a),
                                                      omega, t_ref))

    def set_skin(self, skin: float):
        self._check(self.L.vg_set_skin(self.h, float(skin)))

    def set_skin_auto(self, floor: float, cap: float):
        """Skin per scene from its speed at every rebuild (vg_set_skin_auto)."""
        self._check(self.L.vg_set_skin_auto(self.h, float(floor), float(cap)))

    def set_precision(self, mode: int):
        self._check(self.L.vg_set_precision(self.h, int(mode)))

    def set_graphs(self, on: bool):
        """Step whole rebuild periods as captured HIP graphs (default) or
        launch every step from the host (vg_set_graphs)."""
        self._check(self.L.vg_set_graphs(self.h, 1 if on else 0))

    def graph_replays(self) -> int:
        n = C.c_longlong()
        self._check(self.L.vg_graph_replays(self.h, C.byref(n)))
        return n.value

    def info(self):
        sp, on, q, st = C.c_longlong(), C.c_int(), C.c_int(), C.c_ulonglong()
        self._check(self.L.vg_batch_info(self.h, C.byref(sp), C.byref(on), C.byref(q), C.byref(st)))
        return dict(spheres=sp.value, on_device=on.value, queued=q.value, steps=st.value)

    def check_lists(self):
        """vg_check_lists: violations by kind and the first one."""
        return _check_lists(self)

    def stats(self):
        rb, rl, mf, cap, cs = C.c_longlong(), C.c_longlong(), C.c_int(), C.c_int(), C.c_longlong()
        self._check(self.L.vg_stats(self.h, C.byref(rb), C.byref(rl), C.byref(mf), C.byref(cap), C.byref(cs)))
        return dict(rebuilds=rb.value, relists=rl.value, max_fill=mf.value, capacity=cap.value,
                    contact_steps=cs.value)
