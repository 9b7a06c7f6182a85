"""FP32/FP64-mixed pair solves against the FP64 path on identical input
states: the active sets and ImpactStatus flags are identical (the contact
depth stays FP64), and the per-contact and per-sphere impulse errors are
within 1e-5 relative (per-sphere: |I_mix - I_64| / max(|I_64|, 1e-3 max|I_64|);
per-contact likewise for the impulse to the master); in compressed static
beds the per-sphere error is bounded by 1e-4 at the 99th percentile and 1e-3
at most (see _check).  The FP64 path itself
reproduces the reference bit for bit (test_gpu_parity.py), so the errors
are against the FP64 oracle."""
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
    sim.close()
    return dataclasses.replace(wa, pos=st["pos"], vel=st["vel"], ox=st["ox"], oz=st["oz"], av=st["av"],
                               t_current=t)


def _check(r, min_pairs, piles=False):
    """piles: a compressed static bed, where a sphere's net impulse can be a
    thousandth of its balanced contact impulses, so the FP32 rounding of the
    contact forces (a few 1e-7 of each) exceeds 1e-5 of the net; measured
    there: per-sphere 99th percentile about 6e-5, maximum about 1e-4.  The
    per-contact bound 1e-5 holds everywhere."""
    assert r["active_pairs"] >= min_pairs, r
    assert r["active_sets_identical"], r
    assert r["status_flags_identical"], r
    assert r["contact_err_max"] <= TOL, r
    if piles:
        assert r["sphere_impulse_err_p99"] <= 1e-4, r
        assert r["sphere_impulse_err_max"] <= 1e-3, r
        assert r["sphere_ang_impulse_err_max"] <= 1e-3, r
    else:
        assert r["sphere_impulse_err_max"] <= TOL, r
        assert r["sphere_ang_impulse_err_max"] <= TOL, r


@pytest.mark.parametrize("name", ["box_bsdbf", "box_h", "box_hertz", "box_btsdbf", "box_b2f", "drum_kin"])
def test_mixed_step_matches_fp64(name):
    from vedo_amd.precision import compare_precisions
    z, wa = load_golden(name)
    wa2 = _advanced(wa)
    for k in (1, 5):
        _check(compare_precisions(wa2, steps=k), 5, piles=name == "drum_kin")


def test_mixed_step_matches_fp64_packed_drum():
    # 20,000 spheres of the benchmark's packed bed (pre-compressed contacts)
    from vedo_amd import scenes
    from vedo_amd.precision import compare_precisions
    wa, _, _, _ = scenes.drum_scene(20000, arrays=True, up_iact=20)
    r = compare_precisions(wa, steps=1)
    _check(r, 10000, piles=True)


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
