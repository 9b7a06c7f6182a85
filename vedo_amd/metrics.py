"""Long-run statistics of a scene's state: wall pressure, kinetic energy and
packing fraction, the quantities chaotic runs are compared on.

Wall pressure has no function in the reference; it is formed from the
reference's own per-contact output, the ImpactStatus impulse to the master
(ImpactStatus::SetImpactInformation, ISwBSDBF.cpp:133, ISwH.cpp, ISwHertz.cpp):
for each wall, the sum of |ImpactToMaster| over its active contacts of the
last step, divided by the step's dt and the wall's area (quasi-plate:
width x length of its mid-plane rectangle, less the hole of a plate with a
circular hole; quasi-cylinder: its lateral area 2 pi r h).  The records are those vg_get_active exports (IactRecordTab order
is irrelevant: they are summed in (lo, hi) key order, sequentially, so that
equal records give equal bits)."""
from __future__ import annotations

import math

import numpy as np

from .world import QUASICYLINDER, QUASIPLATE, QUASIPLATE_HOLE, SPHERE


def wall_areas(wa, ids=None):
    """Area of each DO's wall (NaN for spheres), in DO-id order (or for the
    DOs `ids` of the scene's original numbering)."""
    model = wa.model_of if ids is None else wa.model_of[np.asarray(ids)]
    out = np.full(len(model), np.nan)
    for k, m in enumerate(model):
        shape = int(wa.model_shape[m])
        a = wa.model_attr[m]
        if shape == QUASIPLATE:
            out[k] = a[0] * a[2]
        elif shape == QUASIPLATE_HOLE:
            # the hole is cut out where it lies inside the plate (as the
            # volume of DOModel::Analysis, DOModel.cpp:108-116)
            r, hx, hy = wa.model_hole[m]
            inside = abs(hx) <= 0.5 * a[0] - r and abs(hy) <= 0.5 * a[2] - r
            out[k] = a[0] * a[2] - (math.pi * r * r if inside else 0.0)
        elif shape == QUASICYLINDER:
            out[k] = 2.0 * math.pi * a[0] * a[1]
    return out


def wall_pressure(lo, hi, status, wa, dt, ids=None):
    """{wall DO id: pressure} from the active records (lo, hi, statuses) of
    the last step; `ids` maps the current DOs to the description's (after a
    zone of interest erased some), None when none were erased."""
    lo = np.asarray(lo)
    hi = np.asarray(hi)
    order = np.lexsort((hi, lo))
    model = wa.model_of if ids is None else wa.model_of[np.asarray(ids)]
    area = wall_areas(wa, ids)
    walls = np.nonzero(wa.model_shape[model] != SPHERE)[0]
    sums = {int(w): 0.0 for w in walls}
    tm = np.asarray(status["to_master"], np.float64)
    for q in order:
        w = int(hi[q])
        if w in sums:
            v = tm[q]
            sums[w] += math.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
    return {w: s / (dt * area[w]) for w, s in sums.items()}


def kinetic_energy(state, wa, ids=None):
    """Translational and rotational kinetic energy of the spheres (masses and
    moments of inertia of DOModel::Analysis, DOModel.cpp:17-149)."""
    model = wa.model_of if ids is None else wa.model_of[np.asarray(ids)]
    sph = wa.model_shape[model] == SPHERE
    r = wa.model_attr[model[sph], 0]
    m = wa.model_density[model[sph]] * (4.0 / 3.0 * math.pi) * r ** 3
    v = state["vel"][sph]
    w = state["av"][sph]
    return (0.5 * float(np.sum(m * np.einsum("ij,ij->i", v, v))),
            0.5 * float(np.sum(0.4 * m * r * r * np.einsum("ij,ij->i", w, w))))
