"""The reference's .ido checkpoint format: read, write, and the restart of a
GPU scene from it.

An .ido file (DOWorld::WriteIDO / ReadIDO, src/vedo/framework/DOWorld_RWIDO.cpp:9-202)
is the SystemParameter (SystemParameter.cpp:61-136), the DOModels
(DOModel.cpp:267-330), the IactModels (IactModel.cpp:102-140), the DOStatus
list (DOStatus.cpp:176-205) and the IactRecordTab (IactRecordTab.cpp:346-539),
little-endian, strings as a 32-bit length and the bytes
(aries/utility/BinaryIO.cpp:8-21), floats FP64, a Vector3df three FP64, the
DOShapeAttributes union 80 bytes.  This build writes no user-defined values
(NumUDDDOStatus = NumUDDImpactStatus = 0, the reference's default).

read_ido / write_ido round-trip a file byte for byte.  ido_world turns one
into the WorldArrays a GPU scene is created from, and ido_history into the
records vg_set_history restores: as the reference's reader leaves them, with
LastContact and LastBond set (IactRecordTab.cpp:453-504 fills one
ImpactStatus whose default constructor sets both, ImpactStatus.cpp:8-9, and
only ever assigns Contact and Bond).  record_ido writes a GPU scene's state
and records as RecordIDO does.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from .world import (BEHAVIORS, QUASICYLINDER, QUASIPLATE, QUASIPLATE_HOLE, SCOPES, SHAPES, SPHERE, DOModel, DOStatus,
                    DOWorld, IactModel, SystemParameter, VedoError, WorldArrays, record_order)

ATTR_BYTES = 80          # sizeof(DOShapeAttributes)
RELEASE = "2016-09-23"   # aries::information::_ReleaseDate
RECORD_DTYPE = np.dtype([("first", "<u4"), ("second", "<u4"), ("contact", "u1"), ("bond", "u1"),
                         ("kn", "<f8"), ("init_vel", "<f8"), ("shear", "<f8", 3), ("point", "<f8", 3),
                         ("dir", "<f8", 3), ("to_master", "<f8", 3), ("ang_to_master", "<f8", 3),
                         ("overlap", "<f8")], align=False)
assert RECORD_DTYPE.itemsize == 4 + 4 + 1 + 1 + 8 * 18


@dataclass
class IdoModel:
    name: str
    group: str
    behavior: int
    scope: int
    density: float
    density_factor: float
    fext: Tuple[float, float, float]
    color: int
    shape: int
    attr: bytes                      # the DOShapeAttributes union, as stored
    materials: List[Tuple[str, float]] = field(default_factory=list)

    def attributes(self) -> Tuple[float, ...]:
        return struct.unpack_from("<10d", self.attr)


@dataclass
class IdoIact:
    master: str
    slave: str
    equation: str
    mechanisms: List[Tuple[str, float]]


@dataclass
class IdoStatus:
    name: str
    id: int
    vectors: np.ndarray              # [7,3]: position, velocity, Ox, Oz, angular velocity, impact, angular impact


@dataclass
class Ido:
    release: str
    udd_dos: int
    udd_is: int
    max_id: int
    safety: float
    t_start: float
    t_stop: float
    t_current: float
    dt: float
    g: Tuple[float, float, float]
    zoi_switch: bytes
    zoi_lo: Tuple[float, float, float]
    zoi_hi: Tuple[float, float, float]
    pbc_switch: bytes
    pbc_lo: Tuple[float, float, float]
    pbc_hi: Tuple[float, float, float]
    models: List[IdoModel]
    iacts: List[IdoIact]
    statuses: List[IdoStatus]
    records: np.ndarray              # RECORD_DTYPE, in the table's (std::map) key order


class _R:
    def __init__(self, b: bytes):
        self.b, self.o = b, 0

    def take(self, n):
        v = self.b[self.o:self.o + n]
        if len(v) != n:
            raise ValueError("truncated .ido file")
        self.o += n
        return v

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def vec(self):
        return struct.unpack("<3d", self.take(24))

    def string(self):
        return self.take(self.u32()).decode("latin-1")


def read_ido(path: str) -> Ido:
    """DOWorld::ReadIDO(filename, IactRecordTab*) (DOWorld_RWIDO.cpp:9-55)."""
    with open(path, "rb") as f:
        r = _R(f.read())
    release = r.string()
    if r.u32() != 3:
        raise ValueError("SystemParameter: size of integer data is wrong")
    udd_dos, udd_is, max_id = r.u32(), r.u32(), r.u32()
    if r.u32() != 1:
        raise ValueError("SystemParameter: size of float data is wrong")
    safety = r.f64()
    t_start, t_stop, t_current, dt = r.f64(), r.f64(), r.f64(), r.f64()
    g = r.vec()
    zs, zl, zh = r.take(3), r.vec(), r.vec()
    ps, pl, ph = r.take(3), r.vec(), r.vec()
    if udd_dos or udd_is:
        raise ValueError("user-defined values in the .ido file are not supported")
    models = []
    for _ in range(r.u32()):
        name, group = r.string(), r.string()
        beh, scope = r.u32(), r.u32()
        dens, df = r.f64(), r.f64()
        fext = r.vec()
        color, shape = r.u32(), r.u32()
        attr = r.take(ATTR_BYTES)
        mats = [(r.string(), r.f64()) for _ in range(r.u32())]
        if shape not in SHAPES.values():
            raise ValueError(f"DOModel {name}: shape {shape} is not supported")
        models.append(IdoModel(name, group, beh, scope, dens, df, fext, color, shape, attr, mats))
    iacts = []
    for _ in range(r.u32()):
        m, s, eq = r.string(), r.string(), r.string()
        iacts.append(IdoIact(m, s, eq, [(r.string(), r.f64()) for _ in range(r.u32())]))
    statuses = []
    for _ in range(r.u32()):
        name, ident = r.string(), r.u32()
        v = np.frombuffer(r.take(7 * 24), "<f8").reshape(7, 3).copy()
        statuses.append(IdoStatus(name, ident, v))
    n = r.u32()
    recs = np.frombuffer(r.take(n * RECORD_DTYPE.itemsize), RECORD_DTYPE).copy() if n else np.zeros(0, RECORD_DTYPE)
    if r.o != len(r.b):
        raise ValueError("trailing bytes after the IactRecordTab")
    return Ido(release, udd_dos, udd_is, max_id, safety, t_start, t_stop, t_current, dt, g, zs, zl, zh, ps, pl, ph,
               models, iacts, statuses, recs)


def _ws(out: list, s: str):
    b = s.encode("latin-1")
    out.append(struct.pack("<I", len(b)) + b)


def write_ido(path: str, d: Ido):
    """DOWorld::WriteIDO(filename, IactRecordTab*) (DOWorld_RWIDO.cpp:149-202)."""
    out: List[bytes] = []
    _ws(out, d.release)
    out.append(struct.pack("<IIIII", 3, d.udd_dos, d.udd_is, d.max_id, 1))
    out.append(struct.pack("<d", d.safety))
    out.append(struct.pack("<4d", d.t_start, d.t_stop, d.t_current, d.dt))
    out.append(struct.pack("<3d", *d.g))
    out.append(bytes(d.zoi_switch) + struct.pack("<6d", *d.zoi_lo, *d.zoi_hi))
    out.append(bytes(d.pbc_switch) + struct.pack("<6d", *d.pbc_lo, *d.pbc_hi))
    out.append(struct.pack("<I", len(d.models)))
    for m in d.models:
        _ws(out, m.name)
        _ws(out, m.group)
        out.append(struct.pack("<II2d3dII", m.behavior, m.scope, m.density, m.density_factor, *m.fext, m.color,
                               m.shape))
        assert len(m.attr) == ATTR_BYTES
        out.append(bytes(m.attr))
        out.append(struct.pack("<I", len(m.materials)))
        for k, v in m.materials:
            _ws(out, k)
            out.append(struct.pack("<d", v))
    out.append(struct.pack("<I", len(d.iacts)))
    for it in d.iacts:
        _ws(out, it.master)
        _ws(out, it.slave)
        _ws(out, it.equation)
        out.append(struct.pack("<I", len(it.mechanisms)))
        for k, v in it.mechanisms:
            _ws(out, k)
            out.append(struct.pack("<d", v))
    out.append(struct.pack("<I", len(d.statuses)))
    for st in d.statuses:
        _ws(out, st.name)
        out.append(struct.pack("<I", st.id))
        out.append(np.ascontiguousarray(st.vectors, "<f8").tobytes())
    out.append(struct.pack("<I", len(d.records)))
    out.append(np.ascontiguousarray(d.records, RECORD_DTYPE).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))


def ido_world(d: Ido, up_iact: int, record: int) -> WorldArrays:
    """The DOWorld of an .ido (DOWorld::ReadIDO, DOWorld_RWIDO.cpp:9-55) as the
    WorldArrays a scene is created from; the SystemParameter's safety factor
    becomes the scene's, as SystemParameter::operator<< sets it
    (SystemParameter.cpp:116-117).  The zone of interest and the periodic
    boundary are taken from their lower and upper points, whose differences
    decide the switches (Boundary::Correct after reading,
    SystemParameter.cpp:130-139)."""
    shape_name = {v: k for k, v in SHAPES.items()}
    beh_name = {v: k for k, v in BEHAVIORS.items()}
    scope_name = {v: k for k, v in SCOPES.items()}
    nattr = {SPHERE: 1, QUASICYLINDER: 2, QUASIPLATE: 3, QUASIPLATE_HOLE: 6}
    models = [DOModel(m.name, m.group, beh_name[m.behavior], scope_name[m.scope], m.density, m.density_factor,
                      tuple(m.fext), shape_name[m.shape], m.attributes()[:nattr[m.shape]]) for m in d.models]
    iacts = [IactModel(i.master, i.slave, i.equation, dict(i.mechanisms)) for i in d.iacts]
    sts = [DOStatus(s.name, *(tuple(v) for v in s.vectors[:5])) for s in d.statuses]
    sp = SystemParameter(d.t_start, d.t_stop, d.dt, d.t_current, tuple(d.g), d.safety, tuple(d.zoi_lo),
                         tuple(d.zoi_hi), tuple(d.pbc_lo), tuple(d.pbc_hi))
    return DOWorld(sp, models, iacts, sts).to_arrays(up_iact, record)


def ido_history(d: Ido):
    """(lo, hi, statuses) for vg_set_history / OracleSim.set_history, with the
    reader's LastContact = LastBond = true."""
    from .gpu import STATUS_DTYPE
    r = d.records
    st = np.zeros(len(r), STATUS_DTYPE)
    st["contact"] = r["contact"]
    st["bond"] = r["bond"]
    st["last_contact"] = 1
    st["last_bond"] = 1
    for f in ("kn", "init_vel", "shear", "point", "dir", "to_master", "ang_to_master", "overlap"):
        st[f] = r[f]
    return r["first"].copy(), r["second"].copy(), st


def record_ido(path: str, wa: WorldArrays, state, time: float, lo, hi, st, template: Ido = None,
               model_of=None):
    """Write a scene's state and IactRecordTab as RecordIDO does
    (Consultant.cpp:213-229): the models and IactModels of the description
    (or of `template`, an .ido the scene was read from), the DOStatus list of
    `state` (a GpuSimulation.state() dict) and the records (lo, hi, statuses
    as vg_get_active returns them, in key order).  Without `template` the
    models and IactModels come from `wa` itself.

    `model_of` is the model index of each DO of `state` (GpuSimulation.models()).
    It may be left out only while the scene still holds every DO of `wa`: once
    a zone of interest has erased DOs (Consultant::CleanUp), the remaining ones
    are renumbered and `wa.model_of` no longer names them."""
    names = wa.model_names
    n = len(state["pos"])
    if model_of is None:
        if n != wa.n_do:
            raise VedoError(f"the state holds {n} DOs, the description {wa.n_do}: pass model_of "
                            "(the model of each remaining DO)")
        model_of = wa.model_of
    model_of = np.asarray(model_of)
    if len(model_of) != n:
        raise VedoError(f"model_of names {len(model_of)} DOs, the state holds {n}")
    if template is not None:
        models, iacts = template.models, template.iacts
        head = dict(release=template.release, max_id=template.max_id, zoi_switch=template.zoi_switch,
                    zoi_lo=template.zoi_lo, zoi_hi=template.zoi_hi, pbc_switch=template.pbc_switch,
                    pbc_lo=template.pbc_lo, pbc_hi=template.pbc_hi, t_stop=template.t_stop)
    else:
        # from the description alone: colour 0, no material options, the
        # shape union holding the model's sizes (zero elsewhere), one
        # IactModel per group pair with a law (master the lower group)
        from .xmlio import _iact_models
        models = []
        for k, name in enumerate(names):
            a = list(wa.model_attr[k]) + list(wa.model_hole[k]) + [0.0] * 4
            models.append(IdoModel(name, wa.group_names[int(wa.model_group[k])], int(wa.model_behavior[k]),
                                   int(wa.model_scope[k]), float(wa.model_density[k]),
                                   float(wa.model_density_factor[k]), tuple(float(x) for x in wa.model_fext[k]), 0,
                                   int(wa.model_shape[k]), struct.pack("<10d", *a)))
        iacts = [IdoIact(m, s, eq, mech) for m, s, eq, mech in _iact_models(wa)]
        lo3 = np.minimum(wa.zoi_lo, wa.zoi_hi)
        hi3 = np.maximum(wa.zoi_lo, wa.zoi_hi)
        plo = np.minimum(wa.pbc_lo, wa.pbc_hi)
        phi = np.maximum(wa.pbc_lo, wa.pbc_hi)
        head = dict(release=RELEASE, max_id=0, zoi_switch=bytes(int(x) for x in (lo3 != hi3)),
                    zoi_lo=tuple(float(x) for x in lo3), zoi_hi=tuple(float(x) for x in hi3),
                    pbc_switch=bytes(int(x) for x in (plo != phi)), pbc_lo=tuple(float(x) for x in plo),
                    pbc_hi=tuple(float(x) for x in phi), t_stop=wa.t_stop)
    vectors = np.stack([state[k] for k in ("pos", "vel", "ox", "oz", "av", "impact", "ang_impact")], 1)
    sts = [IdoStatus(names[int(model_of[i])], 0, vectors[i]) for i in range(n)]
    order = record_order(lo, hi)
    recs = np.zeros(len(order), RECORD_DTYPE)
    recs["first"], recs["second"] = np.asarray(lo)[order], np.asarray(hi)[order]
    for f in ("contact", "bond", "kn", "init_vel", "shear", "point", "dir", "to_master", "ang_to_master", "overlap"):
        recs[f] = st[f][order]
    d = Ido(head["release"], 0, 0, head["max_id"], wa.safety_factor, wa.t_start, head["t_stop"], time, wa.dt,
            tuple(float(x) for x in wa.g), head["zoi_switch"], head["zoi_lo"], head["zoi_hi"], head["pbc_switch"],
            head["pbc_lo"], head["pbc_hi"], models, iacts, sts, recs)
    write_ido(path, d)
    return d
