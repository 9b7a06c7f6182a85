"""FP32/FP64-mixed pair solves against FP64 on identical input states: the
FP64 path of this library (which reproduces the reference bit for bit,
test_gpu_parity.py) from advanced states with their contact histories, and
the reference's own dumps (tests/golden) for the first step of every
fixture.  Bars for one step from identical states, everywhere, compressed
beds included: identical active sets and ImpactStatus flags (the contact
depth stays FP64); per-contact impulse error <= 1e-5 relative
(|F_mix - F_64| / max(|F_64|, 1e-3 max |F_64|)); per-sphere impulse and
angular impulse error <= 1e-5 (|I_mix - I_64| / max(|I_64|, 1e-3 max_j
|I_64,j|), the angular one floored at 1e-3 x median radius x largest pair
impulse).  After five steps, when the two trajectories have drifted apart,
the same errors stay below 1e-4."""
import dataclasses

import numpy as np
import pytest

from conftest import load_golden

TOL = 1e-5

pytestmark = pytest.mark.gpu


def _advanced(wa, min_active=5, chunk=20, max_steps=4000):
    """The DOWorld after enough FP64 steps for `min_active` contacts, as WorldArrays."""
    from vedo_amd.gpu import GpuSimulation
    sim = GpuSimulation(wa)
    done = 0
    while done < max_steps and sim.count_active() < min_active:
        sim.step(chunk)
        done += chunk
    st = sim.state()
    t = sim.time
    hist = sim.active()
    sim.close()
    return dataclasses.replace(wa, pos=st["pos"], vel=st["vel"], ox=st["ox"], oz=st["oz"], av=st["av"],
                               t_current=t), hist


def _check(r, min_pairs, tol=TOL):
    print(f"\n  {r}", flush=True)
    assert r["active_pairs"] >= min_pairs, r
    assert r["active_sets_identical"], r
    assert r["status_flags_identical"], r
    assert r["contact_err_max"] <= tol, r
    assert r["sphere_impulse_err_max"] <= tol, r
    assert r["sphere_ang_impulse_err_max"] <= tol, r


@pytest.mark.parametrize("name", ["box_bsdbf", "box_h", "box_hertz", "box_btsdbf", "box_b2f", "drum_kin", "box_crowd"])
def test_mixed_step_matches_fp64(name):
    from vedo_amd.precision import compare_precisions
    z, wa = load_golden(name)
    wa2, hist = _advanced(wa)
    # one step from identical states: the 1e-5 bar; five steps: the two
    # trajectories have drifted apart for four steps (bar 1e-4)
    _check(compare_precisions(wa2, steps=1, history=hist), 5)
    _check(compare_precisions(wa2, steps=5, history=hist), 5, tol=1e-4)


def test_mixed_step_matches_fp64_packed_drum():
    # 20,000 spheres of the benchmark's packed bed (pre-compressed contacts)
    from vedo_amd import scenes
    from vedo_amd.precision import compare_precisions
    wa, _, _, _ = scenes.drum_scene(20000, arrays=True, up_iact=20)
    r = compare_precisions(wa, steps=1)
    _check(r, 10000)


def test_precision_switch_is_validated():
    from vedo_amd.gpu import PRECISION_FP64, PRECISION_MIXED, GpuError, GpuSimulation
    z, wa = load_golden("box_bsdbf")
    sim = GpuSimulation(wa)
    assert sim.precision == PRECISION_FP64
    sim.set_precision(PRECISION_MIXED)
    assert sim.precision == PRECISION_MIXED
    with pytest.raises(GpuError):
        sim.set_precision(7)
    sim.close()


@pytest.mark.parametrize("name", ["box_bsdbf", "box_h", "box_hertz", "box_btsdbf", "box_b2f", "box_bond",
                                  "box_bond_bt", "box_cyl", "box_orbital", "box_piston", "box_hole", "box_pbc_xy",
                                  "box_pbc_xyz", "box_pbc_x", "box_pbc_h", "drum_kin"])
def test_mixed_first_step_matches_reference_dump(name):
    """One mixed step from the fixture's start against the reference's own
    impulses of that step (dump 1) and its active set."""
    from vedo_amd.gpu import PRECISION_MIXED, GpuSimulation
    from vedo_amd.precision import _rel_err
    from conftest import step_golden
    z, wa = load_golden(name)
    assert int(z["dumps"][1]) == 1
    sim = GpuSimulation(wa)
    sim.set_precision(PRECISION_MIXED)
    step_golden(sim, z, 0)
    st = sim.state()
    ref = z["d1_state"]
    sph = wa.model_shape[wa.model_of] == 1
    ri = z["d1_iact"]
    act = ri["active"].astype(bool)
    e_lin = _rel_err(st["impact"][sph], ref[sph, 5])
    # angular impulses floored by the torque the contact impulses can exert
    # (median radius x largest pair impulse), as precision.compare_precisions
    f_scale = float(np.linalg.norm(z["d1_status"]["to_master"][act], axis=1).max())
    r_med = float(np.median(wa.model_attr[wa.model_of[sph], 0]))
    e_ang = _rel_err(st["ang_impact"][sph], ref[sph, 6], scale=r_med * f_scale)
    rset = sorted(zip(np.minimum(ri["slave"], ri["master"])[act].tolist(),
                      np.maximum(ri["slave"], ri["master"])[act].tolist()))
    lo, hi, _ = sim.active()
    sim.close()
    print(f"\n  [{name}] per-sphere impulse error {e_lin.max():.2e}, angular {e_ang.max():.2e}, "
          f"active {len(lo)}/{len(rset)}", flush=True)
    assert sorted(zip(lo.tolist(), hi.tolist())) == rset
    assert e_lin.max() <= TOL and e_ang.max() <= TOL
