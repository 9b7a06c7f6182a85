"""The reference's DOWorld XML scene file, read and written without Boost.

write_xml follows DOWorld::WriteXML(filename, IactRecordTab*)
(DOWorld_WriteXML.cpp:19-498) character for character: iostream's default
float format (printf %g, 6 significant digits) in the parameter and model
sections, %.6e for positions, velocities, angular velocities and impulses,
%.6f for orientations (the stream flags it sets around each DOStatus and
IactStatus), booleans as 0/1.  read_xml follows DOWorld::ReadXML
(DOWorld_ReadXML.cpp:216-695): the defaults it applies (DensityFactor 1,
Color bylayer, Scope "undefined" -> local for mobile models and global
otherwise, no ZOI/PBC, zero impulses; ZOI and PBC bounds default to 0
per missing attribute), the in-place normalisation of both
orientations (Vector3d::normalized, Vector3d.h:128-141), the SimConstant
named "SafetyFactor" clamped to at least 1 (the writer names it
"ContactDetectSafetyFactor", which the reader does not look for: a file the
reference wrote reads back with the default 1.1), and ImpactStatus records
with LastContact = Contact and LastBond = Bond (the constructor at
ImpactStatus.cpp:86-95).

Only what the GPU path runs is accepted: spheres, quasi-plates (also with a
circular hole) and quasi-cylinders; other shapes and user-defined values are
rejected with VedoError.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from typing import List, Optional, Tuple

import numpy as np

from .world import (BEHAVIORS, LAWS, QUASICYLINDER, QUASIPLATE, QUASIPLATE_HOLE, SCOPES, SHAPES, SPHERE, DOModel,
                    DOStatus, DOWorld, IactModel, SystemParameter, VedoError, WorldArrays, record_order)

VERSION = "X6 build 923"   # aries::information::_Version (Constants.h:44)
DEFAULT_SAFETY = 1.1       # aries::Constants' initial SafetyFactor (Constants.h:184)


def _g(x: float) -> str:
    return "%g" % x


def _e(x: float) -> str:
    return "%.6e" % x


def _f(x: float) -> str:
    return "%.6f" % x


def _vec(tag: str, v, fmt, pad: str) -> str:
    return f'{pad}<{tag}x="{fmt(v[0])}" y="{fmt(v[1])}" z="{fmt(v[2])}"/>'


def _iact_models(wa: WorldArrays):
    """The IactModels of a WorldArrays' dense group table: one per group pair
    (i <= j) that has a law, master group i, mechanisms in the order the
    solver's constructor reads them."""
    law_name = {code: name for name, (code, _) in LAWS.items()}
    g = len(wa.group_names)
    out = []
    for i in range(g):
        for j in range(i, g):
            code = int(wa.iact_law[i, j])
            if code:
                names = LAWS[law_name[code]][1]
                out.append((wa.group_names[i], wa.group_names[j], law_name[code],
                            [(n, float(wa.iact_param[i, j, k])) for k, n in enumerate(names)]))
    return out


def write_xml(path: str, wa: WorldArrays, state: Optional[dict] = None, records=None, time: Optional[float] = None,
              max_id: int = 0):
    """DOWorld::WriteXML of a scene: its models and IactModels, the DOStatus
    list of `state` (a state() dict; the scene's initial statuses with zero
    impulses when None) and, when given, the records (lo, hi, statuses in
    the layout active() returns) as the IactStatusTab in key order."""
    shape_name = {v: k for k, v in SHAPES.items()}
    beh_name = {v: k for k, v in BEHAVIORS.items()}
    scope_name = {v: k for k, v in SCOPES.items()}
    if state is None:
        z = np.zeros_like(wa.pos)
        state = dict(pos=wa.pos, vel=wa.vel, ox=wa.ox, oz=wa.oz, av=wa.av, impact=z, ang_impact=z)
    t_cur = wa.t_current if time is None else time
    L = ['<?xml version="1.0" encoding="UTF-8"?>',
         '<DOWorld xmlns:xsi="http://www.w3.org/2001/XMLSchema-instance" '
         'xsi:noNamespaceSchemaLocation="DOWorldSchema.xsd">',
         f"\t<Version>{VERSION}</Version>", "\t<SimParameter>", "\t\t<SimConstant>",
         f'\t\t\t<Constant Name="ContactDetectSafetyFactor" Value="{_g(wa.safety_factor)}"/>',
         '\t\t\t<Constant Name="NumUDDDOStatus" Value="0"/>',
         '\t\t\t<Constant Name="uNumUDDImpactStatus" Value="0"/>',
         f'\t\t\t<Constant Name="MaxIDofDOStatus" Value="{int(max_id)}"/>', "\t\t</SimConstant>",
         f'\t\t<TimeControl Start="{_g(wa.t_start)}" Stop="{_g(wa.t_stop)}" Interval="{_g(wa.dt)}" '
         f'Current="{_g(t_cur)}"/>',
         f'\t\t<FieldAcceleration x="{_g(wa.g[0])}" y="{_g(wa.g[1])}" z="{_g(wa.g[2])}"/>']
    lo = np.minimum(wa.zoi_lo, wa.zoi_hi)
    hi = np.maximum(wa.zoi_lo, wa.zoi_hi)
    on = lo != hi
    if on.any():
        s = "\t\t<ZOI"
        for c, a in enumerate("XYZ"):
            if on[c]:
                s += f' {a}Min="{_g(lo[c])}" {a}Max="{_g(hi[c])}"'
        L.append(s + "/>")
    # the periodic boundary the same way (DOWorld_WriteXML.cpp:76-95)
    lo = np.minimum(wa.pbc_lo, wa.pbc_hi)
    hi = np.maximum(wa.pbc_lo, wa.pbc_hi)
    on = lo != hi
    if on.any():
        s = "\t\t<PBC"
        for c, a in enumerate("XYZ"):
            if on[c]:
                s += f' {a}Min="{_g(lo[c])}" {a}Max="{_g(hi[c])}"'
        L.append(s + "/>")
    L += ["\t</SimParameter>", "\t<DOModelTab>"]
    for k, name in enumerate(wa.model_names):
        shape = int(wa.model_shape[k])
        a, hole = wa.model_attr[k], wa.model_hole[k]
        L.append(f'\t\t<DOModel DOName="{name}" DOGroup="{wa.group_names[int(wa.model_group[k])]}" '
                 f'Behavior="{beh_name[int(wa.model_behavior[k])]}" Scope="{scope_name[int(wa.model_scope[k])]}" '
                 f'Color="bylayer" Density="{_g(wa.model_density[k])}" '
                 f'DensityFactor="{_g(wa.model_density_factor[k])}">')
        L.append(_vec("ExternalForce ", wa.model_fext[k], _g, "\t\t\t"))
        if shape == SPHERE:
            sh = f'<Sphere Radius="{_g(a[0])}"/>'
        elif shape == QUASIPLATE:
            sh = f'<QuasiPlate Width="{_g(a[0])}" Length="{_g(a[2])}" Height="{_g(a[1])}"/>'
        elif shape == QUASIPLATE_HOLE:
            sh = (f'<QuasiPlateWithCircularHole Width="{_g(a[0])}" Length="{_g(a[2])}" Height="{_g(a[1])}" '
                  f'HoleRadius="{_g(hole[0])}" HoleXOffset="{_g(hole[1])}" HoleYOffset="{_g(hole[2])}"/>')
        elif shape == QUASICYLINDER:
            sh = f'<QuasiCylinder Radius="{_g(a[0])}" Height="{_g(a[1])}"/>'
        else:
            raise VedoError(f"shape {shape_name.get(shape, shape)} is not written")
        L += [f"\t\t\t<Shape>{sh}</Shape>", "\t\t</DOModel>"]
    L += ["\t</DOModelTab>", "\t<IactModelTab>"]
    for m, s, eq, mech in _iact_models(wa):
        L.append(f'\t\t<IactModel MasterGroup="{m}" SlaveGroup="{s}" EquationType="{eq}">')
        L += [f'\t\t\t<Mechanism Name="{n}" Value="{_g(v)}"/>' for n, v in mech]
        L.append("\t\t</IactModel>")
    L += ["\t</IactModelTab>", "\t<DOStatusTab>"]
    for i in range(len(state["pos"])):
        L.append(f'\t\t<DOStatus DOName="{wa.model_names[int(wa.model_of[i])]}">')
        L.append(_vec("Position        ", state["pos"][i], _e, "\t\t\t"))
        L.append(_vec("Velocity        ", state["vel"][i], _e, "\t\t\t"))
        L.append(_vec("OrientationX    ", state["ox"][i], _f, "\t\t\t"))
        L.append(_vec("OrientationZ    ", state["oz"][i], _f, "\t\t\t"))
        L.append(_vec("AngularVelocity ", state["av"][i], _e, "\t\t\t"))
        L.append(_vec("Impact          ", state["impact"][i], _e, "\t\t\t"))
        L.append(_vec("AngularImpact   ", state["ang_impact"][i], _e, "\t\t\t"))
        L.append("\t\t</DOStatus>")
    L.append("\t</DOStatusTab>")
    if records is not None and len(records[0]):
        rlo, rhi, st = (np.asarray(x) for x in records)
        L.append("\t<IactStatusTab>")
        for q in record_order(rlo, rhi):
            r = st[q]
            L.append(f'\t\t<IactStatus MasterDOStatusSN="{int(rlo[q])}" SlaveDOStatusSN="{int(rhi[q])}" '
                     f'Contact="{int(bool(r["contact"]))}" Bond="{int(bool(r["bond"]))}"')
            L.append(f'\t\t            NormalStiffness="{_e(r["kn"])}" InitialVelocity="{_e(r["init_vel"])}" '
                     f'Overlap="{_e(r["overlap"])}">')
            L.append(_vec("ShearForce            ", r["shear"], _e, "\t\t\t"))
            L.append(_vec("ImpactPoint           ", r["point"], _e, "\t\t\t"))
            L.append(_vec("ImpactDirection       ", r["dir"], _f, "\t\t\t"))
            L.append(_vec("ImpactToMaster        ", r["to_master"], _e, "\t\t\t"))
            L.append(_vec("AngularImpactToMaster ", r["ang_to_master"], _e, "\t\t\t"))
            L.append("\t\t</IactStatus>")
        L.append("\t</IactStatusTab>")
    L.append("</DOWorld>")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def _v3(el, tag=None):
    e = el if tag is None else el.find(tag)
    return (float(e.get("x")), float(e.get("y")), float(e.get("z")))


def _req(el, tag):
    e = el.find(tag)
    if e is None:
        raise VedoError(f'cannot find tag "{tag}"')
    return e


def _normalized(v):
    """Vector3d::normalized: divide by sqrt((x*x + y*y) + z*z) unless zero."""
    n = math.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
    return v if n == 0 else (v[0] / n, v[1] / n, v[2] / n)


def _shape(el) -> Tuple[str, Tuple[float, ...]]:
    def a(e, k):
        return float(e.get(k, 0.0))
    if (e := el.find("Sphere")) is not None:
        r = a(e, "Radius")
        if r <= 0.0:
            raise VedoError("Sphere: Radius must be positive")
        return "Sphere", (r,)
    if (e := el.find("QuasiPlate")) is not None:
        w, l, h = a(e, "Width"), a(e, "Length"), a(e, "Height")
        if w <= 0.0 or l <= 0.0 or h <= 0.0:
            raise VedoError("QuasiPlate: Width, Length and Height must be positive")
        return "QuasiPlate", (w, h, l)
    if (e := el.find("QuasiPlateWithCircularHole")) is not None:
        w, l, h, hr = a(e, "Width"), a(e, "Length"), a(e, "Height"), a(e, "HoleRadius")
        if w <= 0.0 or l <= 0.0 or h <= 0.0 or hr <= 0.0:
            raise VedoError("QuasiPlateWithCircularHole: sizes must be positive")
        return "QuasiPlateWithCircularHole", (w, h, l, hr, a(e, "HoleXOffset"), a(e, "HoleYOffset"))
    if (e := el.find("QuasiCylinder")) is not None:
        r, h = a(e, "Radius"), a(e, "Height")
        if r <= 0.0 or h <= 0.0:
            raise VedoError("QuasiCylinder: Radius and Height must be positive")
        return "QuasiCylinder", (r, h)
    tags = [c.tag for c in el]
    raise VedoError(f"shape {tags} is not on the GPU path")


def read_xml(path: str, up_iact: int, record: int):
    """DOWorld::ReadXML(filename, IactRecordTab*): returns (WorldArrays,
    records, impulses): records is (lo, hi, statuses) for set_history, or
    None when the file has no IactStatusTab; impulses is (impact,
    angular impact), each [N,3], as the DOStatus list holds them."""
    from .gpu import STATUS_DTYPE
    root = ET.parse(path).getroot()
    if root.tag != "DOWorld":
        raise VedoError('cannot find tag "DOWorld"')
    sp_el = _req(root, "SimParameter")
    safety = DEFAULT_SAFETY
    sc = sp_el.find("SimConstant")
    if sc is not None:
        for c in sc:
            if c.get("Name") == "SafetyFactor":
                safety = max(1.0, float(c.get("Value")))
    g = _v3(_req(sp_el, "FieldAcceleration"))
    tc = _req(sp_el, "TimeControl")
    times = [float(tc.get(k)) for k in ("Start", "Stop", "Interval", "Current")]
    zlo, zhi = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
    z = sp_el.find("ZOI")
    if z is not None:
        zlo = tuple(float(z.get(f"{a}Min", 0.0)) for a in "XYZ")
        zhi = tuple(float(z.get(f"{a}Max", 0.0)) for a in "XYZ")
    plo, phi = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
    p = sp_el.find("PBC")
    if p is not None:   # DOWorld_ReadXML.cpp:287-305
        plo = tuple(float(p.get(f"{a}Min", 0.0)) for a in "XYZ")
        phi = tuple(float(p.get(f"{a}Max", 0.0)) for a in "XYZ")
    models: List[DOModel] = []
    for m in _req(root, "DOModelTab"):
        beh = m.get("Behavior")
        if beh not in BEHAVIORS:
            raise VedoError(f"DOModel {m.get('DOName')}: Behavior {beh} is wrong")
        scope = m.get("Scope", "undefined")
        if scope == "undefined":
            scope = "local" if beh == "mobile" else "global"
        elif scope not in SCOPES:
            raise VedoError(f"DOModel {m.get('DOName')}: Scope {scope} is wrong")
        fe = m.find("ExternalForce")
        shape, attrs = _shape(_req(m, "Shape"))
        models.append(DOModel(m.get("DOName"), m.get("DOGroup"), beh, scope, float(m.get("Density")),
                              float(m.get("DensityFactor", 1.0)), _v3(fe) if fe is not None else (0.0, 0.0, 0.0),
                              shape, attrs))
    iacts = [IactModel(i.get("MasterGroup"), i.get("SlaveGroup"), i.get("EquationType"),
                       {c.get("Name"): float(c.get("Value")) for c in i}) for i in _req(root, "IactModelTab")]
    sts, imp = [], []
    for s in _req(root, "DOStatusTab"):
        sts.append(DOStatus(s.get("DOName"), _v3(s, "Position"), _v3(s, "Velocity"),
                            _normalized(_v3(s, "OrientationX")), _normalized(_v3(s, "OrientationZ")),
                            _v3(s, "AngularVelocity")))
        imp.append([_v3(s, t) if s.find(t) is not None else (0.0, 0.0, 0.0) for t in ("Impact", "AngularImpact")])
    imp = np.array(imp, np.float64).reshape(-1, 2, 3)
    impulses = (imp[:, 0].copy(), imp[:, 1].copy())
    sp = SystemParameter(*times, g, safety, zlo, zhi, plo, phi)
    wa = DOWorld(sp, models, iacts, sts).to_arrays(up_iact, record)
    tab = root.find("IactStatusTab")
    if tab is None:
        return wa, None, impulses
    rows = list(tab)
    lo = np.array([int(r.get("MasterDOStatusSN")) for r in rows], np.uint32)
    hi = np.array([int(r.get("SlaveDOStatusSN")) for r in rows], np.uint32)
    st = np.zeros(len(rows), STATUS_DTYPE)
    for q, r in enumerate(rows):
        c, b = int(r.get("Contact")) != 0, int(r.get("Bond")) != 0
        st[q]["contact"] = st[q]["last_contact"] = c
        st[q]["bond"] = st[q]["last_bond"] = b
        st[q]["kn"] = float(r.get("NormalStiffness"))
        st[q]["init_vel"] = float(r.get("InitialVelocity"))
        st[q]["overlap"] = float(r.get("Overlap"))
        for f, tag in (("shear", "ShearForce"), ("point", "ImpactPoint"), ("dir", "ImpactDirection"),
                       ("to_master", "ImpactToMaster"), ("ang_to_master", "AngularImpactToMaster")):
            st[q][f] = _v3(r, tag)
    # IactRecordTab is a std::map: one record per key, the first pushed kept
    keys = lo.astype(np.int64) << 32 | hi
    _, first = np.unique(keys, return_index=True)
    return wa, (lo[first], hi[first], st[first]), impulses
