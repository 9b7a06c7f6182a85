"""FP64 against FP32/FP64-mixed pair solves on identical input states.

The mixed path (vg_set_precision(h, VG_PRECISION_MIXED)) keeps the contact
depth, the ImpactStatus flags, the histories, the impulse sums and the
integration in FP64 and forms each pair's geometry and forces in FP32.  This
module runs one step of each path from the same DOWorld state and reports:

* whether the active sets (ContactInfo.bActive pairs, with their Contact and
  Bond flags) are identical;
* the per-sphere impulse error |I_mix - I_64| / max(|I_64|, tau max_j |I_64,j|)
  (tau = 1e-3), and the same for the angular impulse with the floor
  tau max(max_j |L_64,j|, median radius x largest pair impulse);
* the per-contact error of the impulse to the master,
  |F_mix - F_64| / max(|F_64|, tau max |F_64|).

The FP64 path reproduces the reference's FP64 results bit for bit (see
tests/test_gpu_parity.py), so these are errors against the FP64 oracle.
"""
from __future__ import annotations

import numpy as np

TAU = 1e-3


def _rel_err(x, ref, tau=TAU, scale=0.0):
    nr = np.linalg.norm(ref, axis=1)
    floor = tau * max(nr.max() if len(nr) else 0.0, scale)
    den = np.maximum(nr, floor)
    num = np.linalg.norm(x - ref, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        e = np.where(den > 0, num / den, np.where(num > 0, np.inf, 0.0))
    return e


def compare_precisions(wa, steps: int = 1, sphere_mask=None, history=None):
    """Step `wa` `steps` times with each precision and compare; returns a dict.
    `history` = (lo, hi, statuses) of the state's IactRecordTab (as
    GpuSimulation.active() returns it) restores the contacts' histories on both
    handles, so that the comparison starts from the full state, shear springs
    included."""
    from .gpu import PRECISION_MIXED, GpuSimulation

    out = {}
    sims = [GpuSimulation(wa), GpuSimulation(wa)]
    try:
        if history is not None:
            for s in sims:
                s.set_history(*history)
        sims[1].set_precision(PRECISION_MIXED)
        for s in sims:
            s.step(steps)
        st = [s.state() for s in sims]
        act = [s.active() for s in sims]
    finally:
        for s in sims:
            s.close()
    if sphere_mask is None:
        sphere_mask = wa.model_shape[wa.model_of] == 1
    (lo0, hi0, s0), (lo1, hi1, s1) = act
    e_lin = _rel_err(st[1]["impact"][sphere_mask], st[0]["impact"][sphere_mask])
    # angular impulses: floored by the torque the contact impulses can exert
    # (median radius x largest pair impulse); without shear they are FP64
    # rounding noise of r x F_normal, which is no scale to measure against
    radius = wa.model_attr[wa.model_of[sphere_mask], 0]
    f_scale = float(np.linalg.norm(s0["to_master"], axis=1).max()) if len(s0) else 0.0
    e_ang = _rel_err(st[1]["ang_impact"][sphere_mask], st[0]["ang_impact"][sphere_mask],
                     scale=float(np.median(radius)) * f_scale if len(radius) else 0.0)
    key0 = lo0.astype(np.uint64) << np.uint64(32) | hi0.astype(np.uint64)
    key1 = lo1.astype(np.uint64) << np.uint64(32) | hi1.astype(np.uint64)
    o0, o1 = np.argsort(key0, kind="stable"), np.argsort(key1, kind="stable")
    same_set = len(key0) == len(key1) and bool(np.array_equal(key0[o0], key1[o1]))
    same_flags = same_set and all(np.array_equal(s0[f][o0], s1[f][o1])
                                  for f in ("contact", "last_contact", "bond", "last_bond"))
    out["active_pairs"] = int(len(key0))
    out["active_sets_identical"] = same_set
    out["status_flags_identical"] = bool(same_flags)
    if same_set and len(key0):
        e_pair = _rel_err(s1["to_master"][o1], s0["to_master"][o0])
        out["contact_err_max"] = float(e_pair.max())
        out["contact_err_p99"] = float(np.percentile(e_pair, 99))
    else:
        out["contact_err_max"] = None
        out["contact_err_p99"] = None
    out["sphere_impulse_err_max"] = float(e_lin.max()) if len(e_lin) else 0.0
    out["sphere_impulse_err_p99"] = float(np.percentile(e_lin, 99)) if len(e_lin) else 0.0
    out["sphere_ang_impulse_err_max"] = float(e_ang.max()) if len(e_ang) else 0.0
    out["tau"] = TAU
    out["steps"] = steps
    return out
