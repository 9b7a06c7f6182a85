import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_golden(name):
    """Load a golden fixture as (npz, WorldArrays)."""
    from vedo_amd.world import WorldArrays
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    sc = z["scalars"]
    wa = WorldArrays(
        model_shape=z["model_shape"], model_behavior=z["model_behavior"], model_scope=z["model_scope"],
        model_group=z["model_group"], model_density=z["model_density"],
        model_density_factor=z["model_density_factor"], model_fext=z["model_fext"],
        model_attr=z["model_attr"], iact_law=z["iact_law"], iact_param=z["iact_param"],
        model_of=z["model_of"], pos=z["pos"], vel=z["vel"], ox=z["ox"], oz=z["oz"], av=z["av"],
        dt=float(sc[0]), t_start=float(sc[1]), t_stop=float(sc[2]), t_current=float(sc[3]), g=z["g"],
        safety_factor=float(sc[4]), up_iact=int(z["cadence"][0]), record=int(z["cadence"][1]))
    if "zoi" in z.files:
        wa.zoi_lo, wa.zoi_hi = z["zoi"][:3].copy(), z["zoi"][3:].copy()
    if "pbc" in z.files:
        wa.pbc_lo, wa.pbc_hi = z["pbc"][:3].copy(), z["pbc"][3:].copy()
    if "model_hole" in z.files:
        wa.model_hole = z["model_hole"].copy()
    return z, wa


GOLDEN_CASES = ["box_bsdbf", "box_h", "box_hertz", "box_btsdbf", "box_b2f", "box_bond", "box_bond_bt", "box_cyl",
                "box_orbital", "box_nbs", "drum_kin", "box_assign", "box_piston", "box_cw3", "box_zoi", "box_hole",
                "box_pbc_xy", "box_pbc_xyz", "box_pbc_x", "box_pbc_h", "box_nbs_quirks", "box_pbc_nbs", "box_crowd"]


def step_golden(sim, z, step):
    """Advance `sim` by the fixture's step `step` (0-based): SimMediator::Run()
    or, when the fixture has assigned statuses, Run(map) with the rows of that
    step that are not NaN (a NaN row is not assigned at that step)."""
    if "assigned" not in z.files:
        sim.step(1)
        return
    a = z["assigned"][step]
    keep = ~np.isnan(a[:, 0])
    ids = np.asarray(z["assigned_ids"])[keep]
    a = a[keep]
    sim.step_assigned(ids, a[:, 0:3], a[:, 3:6], a[:, 6:9], a[:, 9:12], a[:, 12:15])
