"""Wall pressure and kinetic energy (vedo_amd.metrics) on the oracle.

* The pressure formed from the oracle's records equals, bit for bit, the one
  formed from the reference's own ImpactStatus records in the golden dumps
  (ImpactToMaster of every active sphere-wall contact).
* Closed form: a sphere resting on the floor plate of a box presses it with
  its weight, so the floor's pressure is m* g / (w l) once it is at rest.
* kinetic_energy equals DOWorld::CalculateSystemEnergy's translation and
  rotation terms of the oracle (same masses, one summation order apart).
"""
import math

import numpy as np
import pytest

from conftest import load_golden, step_golden
from oracle.oracle import OracleSim
from vedo_amd import metrics, scenes

WALL_CASES = ["box_bsdbf", "box_cyl", "box_orbital", "box_piston", "box_hole", "drum_kin"]


def _fixture_records(z, k):
    ri = z[f"d{k}_iact"]
    act = ri["active"].astype(bool)
    lo = np.minimum(ri["slave"], ri["master"])[act]
    hi = np.maximum(ri["slave"], ri["master"])[act]
    return lo, hi, z[f"d{k}_status"][act]


@pytest.mark.parametrize("name", WALL_CASES)
def test_wall_pressure_from_oracle_records_equals_reference_records(name):
    z, wa = load_golden(name)
    sim = OracleSim(wa)
    step, nonzero = 0, 0
    for k, d in enumerate(z["dumps"]):
        while step < d:
            step_golden(sim, z, step)
            step += 1
        if d == 0:
            continue
        ref = metrics.wall_pressure(*_fixture_records(z, k), wa, wa.dt)
        mine = metrics.wall_pressure(*sim.active(), wa, wa.dt)
        assert ref.keys() == mine.keys()
        for w in ref:
            assert np.float64(mine[w]).view(np.uint64) == np.float64(ref[w]).view(np.uint64), (name, d, w)
        nonzero += sum(1 for v in ref.values() if v > 0)
    assert nonzero > 0, "no wall pressed in any dump"


def test_resting_sphere_presses_the_floor_with_its_weight():
    from vedo_amd.world import DOStatus
    w = scenes.box_scene(1, seed=3, law="ISwBSDBF", dt=2e-5, steps=20000, spacing=2.0, jitter=0.0, vmax=0.0,
                         offset=3.0, kn=2e3, e=0.3, mu=0.4, side=0.01)
    # the sphere in the middle of the box, just above the floor's face
    w.statuses[0] = DOStatus("Sphere", (0.005, 0.005, 0.001 + 0.00101))
    wa = w.to_arrays(up_iact=10, record=10 ** 9)
    sim = OracleSim(wa)
    sim.step(6000)
    p = metrics.wall_pressure(*sim.active(), wa, wa.dt)
    floor = [q for q in p if wa.oz[q][2] == 1.0 and wa.pos[q][2] == 0.0][0]
    r = wa.model_attr[wa.model_of[0], 0]
    m = wa.model_density[wa.model_of[0]] * 4.0 / 3.0 * math.pi * r ** 3
    a = wa.model_attr[wa.model_of[floor]]
    expect = m * 9.81 / (a[0] * a[2])
    print(f"\n  floor pressure {p[floor]:.9e} Pa, m g / A = {expect:.9e} Pa", flush=True)
    assert abs(p[floor] - expect) <= 1e-6 * expect
    assert all(v == 0.0 for q, v in p.items() if q != floor)


def test_kinetic_energy_matches_system_energy():
    z, wa = load_golden("box_bsdbf")
    sim = OracleSim(wa)
    sim.step(40)
    et, er = metrics.kinetic_energy(sim.state(), wa)
    e = sim.system_energy()
    assert abs(et - e["translation"]) <= 1e-13 * e["translation"]
    assert abs(er - e["rotation"]) <= 1e-13 * e["rotation"]
