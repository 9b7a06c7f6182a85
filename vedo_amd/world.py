"""Host-side scene model mirroring VEDO's DOWorld data model.

The classes keep the reference's names and argument meaning so that a scene
described for VEDO maps one-to-one onto the GPU path:

* ``SystemParameter``  -- time control, field acceleration
  (SystemParameter.h:17-25); SafetyFactor lives in aries::Constants
  (Constants.h:184) and is carried here.
* ``DOModel``          -- shape, behaviour, scope, density, DensityFactor,
  external force and shape attributes (DOModel.h:132-137).  ``analysis()``
  restates DOModel::Analysis (DOModel.cpp:17-149) for the quantities the step
  reads.
* ``IactModel``        -- (master group, slave group, EquationType, named
  mechanisms) (IactModel.h:103-107).  Mechanism names are the reference's
  exact strings; a missing name is an error, as GetIactMechanism exits
  (IactModel.cpp:69-90).
* ``DOStatus``         -- per-DO position, velocity, orientation Ox/Oz and
  angular velocity (DOStatus.h:19-22).
* ``DOWorld``          -- the container; ``iact_table()`` builds the dense
  group x group pair table with the reference's first-match symmetric lookup
  (DOWorld.cpp:297-302, IactModel.cpp:16-30).

``to_arrays()`` flattens a world into the structure-of-arrays form that both
the C-ABI (include/vedo_gpu.h) and the test oracle take.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np

# DOShapeType / DOBehaviorType / DOScopeType under -D_STD_CPP_11 (DOModel.h:26-28)
SPHERE, QUASICYLINDER, QUASIPLATE, QUASIPLATE_HOLE = 1, 6, 7, 8
CONSTRAINED, FIXED, MOBILE, ORBITAL = 1, 2, 3, 4
GLOBAL, LOCAL = 1, 2

SHAPES = {"Sphere": SPHERE, "QuasiCylinder": QUASICYLINDER, "QuasiPlate": QUASIPLATE,
          "QuasiPlateWithCircularHole": QUASIPLATE_HOLE}
BEHAVIORS = {"constrained": CONSTRAINED, "fixed": FIXED, "mobile": MOBILE, "orbital": ORBITAL}
SCOPES = {"global": GLOBAL, "local": LOCAL}

# Impact solvers registered in ModuleList.h:52-56, with the mechanism names
# their constructors read, in slot order.
LAWS = {
    "ISwBSDBF": (1, ("NormalStiffness", "ShearStiffness", "NormalDampingRatio", "ShearDampingRatio",
                     "NormalBondStrength", "ShearBondStrength", "FrictionCoefficient")),
    "ISwBSDB2F": (2, ("NormalStiffness", "ShearStiffness", "NormalDampingRatio", "ShearDampingRatio",
                      "NormalBondStrength", "ShearBondStrength", "FrictionCoefficient")),
    "ISwBtSDBF": (3, ("NormalStiffness", "ShearStiffness", "NormalDampingRatio", "ShearDampingRatio",
                      "NormalBondStrength", "ShearBondStrength", "FrictionCoefficient")),
    "ISwH": (4, ("HertzNormalStiffness", "EquivalentMass", "ShearStiffnessCoefficient1",
                 "ShearStiffnessCoefficient2", "DampingRatio", "KineticFrictionCoefficient")),
    "ISwHertz": (5, ("NormalStiffness", "NormalDampingCoefficient")),
}
NPARAM = 8


class VedoError(RuntimeError):
    """Raised where the reference prints an error and calls exit()."""


@dataclass
class SystemParameter:
    time_start: float
    time_stop: float
    time_interval: float
    time_current: float
    field_acceleration: Tuple[float, float, float] = (0.0, 0.0, -9.81)
    safety_factor: float = 1.1
    # ZoneOfInterest (SystemParameter.h:275, Boundary.cpp:79-129): an axis is
    # bounded where its lower and upper values differ; mobile DOs outside are
    # erased at candidate-list rebuilds (Consultant::CleanUp)
    zoi_lower: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    zoi_upper: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    # PeriodicBoundaryConditions (SystemParameter.h:276, Boundary.cpp:79-129):
    # an axis is periodic where its lower and upper values differ
    pbc_lower: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    pbc_upper: Tuple[float, float, float] = (0.0, 0.0, 0.0)


@dataclass
class DOModel:
    name: str
    group: str
    behavior: str
    scope: str
    density: float
    density_factor: float
    external_force: Tuple[float, float, float]
    shape: str
    attributes: Tuple[float, ...]   # Sphere: (r,) QuasiCylinder: (r, h) QuasiPlate: (w, h, l)
    #                                 QuasiPlateWithCircularHole: (w, h, l, hole radius, hole x / y offset)

    def analysis(self):
        """DOModel::Analysis (DOModel.cpp:17-149): range, mass, sudo-mass."""
        pi = 4.0 * math.atan(1.0)
        shape = SHAPES[self.shape]
        a = tuple(self.attributes) + (0.0,) * (6 - len(self.attributes))
        if shape == SPHERE:
            r = a[0]
            rng = r
            vol = (4.0 / 3.0 * pi) * r * r * r
        elif shape == QUASICYLINDER:
            r, h = a[0], a[1]
            rng = max(r, 0.5 * h) if self.scope == "local" else r
            vol = pi * r * r * h
        elif shape == QUASIPLATE:
            w, h, l = a[:3]
            rng = 0.5 * max(max(w, h), l) if self.scope == "local" else 0.5 * h
            vol = w * h * l
        elif shape == QUASIPLATE_HOLE:   # DOModel.cpp:90-120
            w, h, l, hr, hxo, hyo = a
            rng = 0.5 * max(max(w, h), l) if self.scope == "local" else 0.5 * h
            if abs(hxo) <= 0.5 * w - hr and abs(hyo) <= 0.5 * l - hr:
                vol = w * h * l - hr * hr * pi * h
            else:
                vol = w * h * l
        else:
            raise VedoError(f"shape {self.shape} is not on the GPU path")
        mass = self.density * vol
        return rng, mass, self.density_factor * mass


@dataclass
class IactModel:
    master_group: str
    slave_group: str
    equation_type: str
    mechanisms: Dict[str, float]

    def params(self) -> Tuple[int, List[float]]:
        if self.equation_type not in LAWS:
            raise VedoError(f"no ImpactSolver matches EquationType {self.equation_type}")
        law, names = LAWS[self.equation_type]
        out = [0.0] * NPARAM
        for k, n in enumerate(names):
            if n not in self.mechanisms:
                raise VedoError(f"IactModel {self.master_group}/{self.slave_group}: mechanism {n} missing")
            out[k] = float(self.mechanisms[n])
        return law, out


@dataclass
class DOStatus:
    name: str
    position: Sequence[float]
    velocity: Sequence[float] = (0.0, 0.0, 0.0)
    orientation_x: Sequence[float] = (1.0, 0.0, 0.0)
    orientation_z: Sequence[float] = (0.0, 0.0, 1.0)
    angular_velocity: Sequence[float] = (0.0, 0.0, 0.0)


@dataclass
class WorldArrays:
    """Structure-of-arrays form of a DOWorld (DO id order)."""
    model_shape: np.ndarray        # int32 [M]
    model_behavior: np.ndarray     # int32 [M]
    model_scope: np.ndarray        # int32 [M]
    model_group: np.ndarray        # int32 [M]
    model_density: np.ndarray      # f64 [M]
    model_density_factor: np.ndarray
    model_fext: np.ndarray         # f64 [M,3]
    model_attr: np.ndarray         # f64 [M,3]
    iact_law: np.ndarray           # int32 [G,G]
    iact_param: np.ndarray         # f64 [G,G,NPARAM]
    model_of: np.ndarray           # int32 [N]
    pos: np.ndarray                # f64 [N,3]
    vel: np.ndarray
    ox: np.ndarray
    oz: np.ndarray
    av: np.ndarray
    dt: float
    t_start: float
    t_stop: float
    t_current: float
    g: np.ndarray
    safety_factor: float
    up_iact: int
    record: int
    group_names: List[str] = field(default_factory=list)
    model_names: List[str] = field(default_factory=list)
    zoi_lo: np.ndarray = field(default_factory=lambda: np.zeros(3))
    zoi_hi: np.ndarray = field(default_factory=lambda: np.zeros(3))
    model_hole: np.ndarray = None
    pbc_lo: np.ndarray = field(default_factory=lambda: np.zeros(3))
    pbc_hi: np.ndarray = field(default_factory=lambda: np.zeros(3))   # f64 [M,3]: hole radius, x / y offset (QuasiPlateWithCircularHole)

    def __post_init__(self):
        if self.model_hole is None:
            self.model_hole = np.zeros((len(self.model_shape), 3))

    @property
    def n_do(self) -> int:
        return int(self.model_of.shape[0])

    def radius_of(self) -> np.ndarray:
        return self.model_attr[self.model_of, 0]


class DOWorld:
    """Scene container (DOWorld.h:23-227)."""

    def __init__(self, sp: SystemParameter, models: List[DOModel], iact_models: List[IactModel],
                 statuses: List[DOStatus]):
        self.system_parameter = sp
        self.models = list(models)
        self.iact_models = list(iact_models)
        self.statuses = list(statuses)

    def check(self):
        """DOWorld::Check (DOWorld.cpp:849-870): every status names a model."""
        names = {m.name for m in self.models}
        for s in self.statuses:
            if s.name not in names:
                raise VedoError(f"DOStatus names unknown DOModel {s.name}")

    def groups(self) -> List[str]:
        out: List[str] = []
        for m in self.models:
            if m.group not in out:
                out.append(m.group)
        return out

    def iact_table(self, groups: List[str]):
        g = len(groups)
        law = np.zeros((g, g), np.int32)
        par = np.zeros((g, g, NPARAM), np.float64)
        for i, gi in enumerate(groups):
            for j, gj in enumerate(groups):
                for im in self.iact_models:   # first match, either order
                    if (im.master_group == gi and im.slave_group == gj) or \
                       (im.master_group == gj and im.slave_group == gi):
                        k, p = im.params()
                        law[i, j] = k
                        par[i, j, :] = p
                        break
        return law, par

    def to_arrays(self, up_iact: int, record: int) -> WorldArrays:
        self.check()
        groups = self.groups()
        midx = {m.name: k for k, m in enumerate(self.models)}
        law, par = self.iact_table(groups)
        ms = self.models

        def attr3(m):
            a = list(m.attributes) + [0.0] * (6 - len(m.attributes))
            return a[:3]

        def hole3(m):
            a = list(m.attributes) + [0.0] * (6 - len(m.attributes))
            return a[3:6] if SHAPES[m.shape] == QUASIPLATE_HOLE else [0.0, 0.0, 0.0]

        sp = self.system_parameter
        model_of = np.array([midx[s.name] for s in self.statuses], np.int32)
        return WorldArrays(
            model_shape=np.array([SHAPES[m.shape] for m in ms], np.int32),
            model_behavior=np.array([BEHAVIORS[m.behavior] for m in ms], np.int32),
            model_scope=np.array([SCOPES[m.scope] for m in ms], np.int32),
            model_group=np.array([groups.index(m.group) for m in ms], np.int32),
            model_density=np.array([m.density for m in ms], np.float64),
            model_density_factor=np.array([m.density_factor for m in ms], np.float64),
            model_fext=np.array([m.external_force for m in ms], np.float64).reshape(-1, 3),
            model_attr=np.array([attr3(m) for m in ms], np.float64).reshape(-1, 3),
            iact_law=law, iact_param=par, model_of=model_of,
            pos=np.array([s.position for s in self.statuses], np.float64).reshape(-1, 3),
            vel=np.array([s.velocity for s in self.statuses], np.float64).reshape(-1, 3),
            ox=np.array([s.orientation_x for s in self.statuses], np.float64).reshape(-1, 3),
            oz=np.array([s.orientation_z for s in self.statuses], np.float64).reshape(-1, 3),
            av=np.array([s.angular_velocity for s in self.statuses], np.float64).reshape(-1, 3),
            dt=sp.time_interval, t_start=sp.time_start, t_stop=sp.time_stop,
            t_current=sp.time_current, g=np.array(sp.field_acceleration, np.float64),
            safety_factor=sp.safety_factor, up_iact=int(up_iact), record=int(record),
            group_names=groups, model_names=[m.name for m in ms],
            zoi_lo=np.array(sp.zoi_lower, np.float64), zoi_hi=np.array(sp.zoi_upper, np.float64),
            pbc_lo=np.array(sp.pbc_lower, np.float64), pbc_hi=np.array(sp.pbc_upper, np.float64),
            model_hole=np.array([hole3(m) for m in ms], np.float64).reshape(-1, 3))


def record_order(lo, hi) -> np.ndarray:
    """Indices of the records (lo, hi) in IactRecordTab key order, each key
    once with its first record: PushRecord is std::map::insert
    (IactRecordTab.cpp:36-39), and a periodic boundary's traversal can list
    a pair twice (its two interactions carry equal histories)."""
    lo, hi = np.asarray(lo), np.asarray(hi)
    order = np.lexsort((hi, lo))   # stable: equal keys keep their input order
    if len(order) > 1:
        klo, khi = lo[order], hi[order]
        first = np.ones(len(order), bool)
        first[1:] = (klo[1:] != klo[:-1]) | (khi[1:] != khi[:-1])
        order = order[first]
    return order


def write_vscene(wa: WorldArrays, path: str, consultant: int = 0):
    """Write the plain-text scene exchange file read by oracle/ref_harness.cpp.

    Doubles are written as C99 hex floats so every bit survives."""
    h = float.hex
    lines = ["VSCENE 1"]
    lines.append("time %s %s %s %s" % (h(wa.t_start), h(wa.t_stop), h(wa.dt), h(wa.t_current)))
    lines.append("gravity %s %s %s" % tuple(h(float(x)) for x in wa.g))
    lines.append("safety %s" % h(wa.safety_factor))
    if np.any(np.asarray(wa.zoi_lo) != np.asarray(wa.zoi_hi)):
        lines.append("zoi %s %s %s %s %s %s" % tuple(h(float(x)) for x in list(wa.zoi_lo) + list(wa.zoi_hi)))
    if np.any(np.asarray(wa.pbc_lo) != np.asarray(wa.pbc_hi)):
        lines.append("pbc %s %s %s %s %s %s" % tuple(h(float(x)) for x in list(wa.pbc_lo) + list(wa.pbc_hi)))
    lines.append("cadence %d %d %d" % (wa.up_iact, wa.record, consultant))
    lines.append("groups %d %s" % (len(wa.group_names), " ".join(wa.group_names)))
    # hole radius and offsets of QuasiPlateWithCircularHole models, before the models
    holes = [k for k in range(len(wa.model_names)) if wa.model_shape[k] == QUASIPLATE_HOLE]
    if holes:
        lines.append("holes %d" % len(holes))
        for k in holes:
            lines.append("%d %s" % (k, " ".join(h(float(x)) for x in wa.model_hole[k])))
    lines.append("models %d" % len(wa.model_names))
    for k, name in enumerate(wa.model_names):
        lines.append("%s %d %d %d %d %s %s %s %s %s %s %s %s" % (
            name, wa.model_shape[k], wa.model_behavior[k], wa.model_scope[k], wa.model_group[k],
            h(float(wa.model_density[k])), h(float(wa.model_density_factor[k])),
            *(h(float(x)) for x in wa.model_fext[k]), *(h(float(x)) for x in wa.model_attr[k])))
    g = len(wa.group_names)
    pairs = [(i, j) for i in range(g) for j in range(i, g) if wa.iact_law[i, j]]
    lines.append("iact %d" % len(pairs))
    for i, j in pairs:
        lines.append("%d %d %d %s" % (i, j, wa.iact_law[i, j],
                                      " ".join(h(float(x)) for x in wa.iact_param[i, j])))
    lines.append("dos %d" % wa.n_do)
    for i in range(wa.n_do):
        vals = np.concatenate([wa.pos[i], wa.vel[i], wa.ox[i], wa.oz[i], wa.av[i]])
        lines.append("%d %s" % (wa.model_of[i], " ".join(h(float(x)) for x in vals)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
