"""The 4,000,000-sphere hopper (bimodal 0.8 / 1.6 mm at 70/30, two DOGroups,
ISwH between every pair of groups and against the walls, dt 2e-6 s, a zone
of interest below which spheres are erased) at its full size, through
properties that hold at any size:

* Consultant::CleanUp at construction erases exactly the spheres the zone
  excludes (counted on the host), and erasing continues at rebuilds as the
  unsupported bed falls through the zone's bound;
* two runs of the same scene are bitwise identical (the step is
  deterministic: fixed per-sphere summation order, no floating-point
  atomics);
* a restart at a rebuild step from the remaining DOStatus (with their
  DONames) and the exported IactRecordTab continues bitwise as the
  uninterrupted run;
* every index the step, wall and rebuild kernels take from the device
  lists is in range (vg_check_lists) every 2,000 steps and at the end;
* one FP32/FP64-mixed step from the end state (histories carried) against
  one FP64 step: identical active sets and flags, per-contact and
  per-sphere errors <= 1e-5 (the configuration's mixed-precision leg).

The zone's lower bound is raised into the funnel, 60 um under the centre of
the sixth lattice layer (layers 2.2 r_large apart, centres jittered by
+-0.05 r_large), so that erasing starts at once (a sphere falls 28 um in the first 1200
steps of 2e-6 s); the configuration's own bound, 2 r under the outlet, is
first crossed after many thousands of steps.  The lattice starts without
contacts: the run goes on until the lowest layers press on the funnel
plates, so that the restart carries contact histories (ISwH's shear force
and Kn)."""
import dataclasses
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4_000_000
UP = 20
R_LARGE = 1.6e-3
Z_LO = R_LARGE + 5.5 * 2.2 * R_LARGE - 0.06e-3
MIN_CONTACTS = 1000
MAX_STEPS = 20000


def _bits(a, b):
    return int((a.view(np.uint64) != b.view(np.uint64)).sum())


def _spheres(sim, wa):
    return int((wa.model_shape[sim.models()] == 1).sum())


def test_hopper_4m_zone_determinism_restart():
    from vedo_amd import scenes
    from vedo_amd.gpu import GpuSimulation
    wa = scenes.hopper_cfg3(N, up_iact=UP)
    wa.zoi_lo = np.array([0.0, 0.0, Z_LO])
    sph = wa.model_shape[wa.model_of] == 1
    expected = int((sph & (wa.pos[:, 2] >= Z_LO)).sum())
    print(f"\n  {N} spheres; zone from z = {Z_LO} m: {int(sph.sum()) - expected} below it at the start", flush=True)
    t0 = time.perf_counter()
    a = GpuSimulation(wa)
    n0 = _spheres(a, wa)
    print(f"  created ({time.perf_counter() - t0:.1f} s): {n0} spheres left", flush=True)
    assert n0 == expected
    assert a.n_do == n0 + int((~sph).sum())
    b = GpuSimulation(wa)
    counts = [n0]
    done = 0
    # the lattice starts without contacts: the lowest layers reach the funnel
    # plates after some thousands of steps
    while done < MAX_STEPS and a.count_active() < MIN_CONTACTS:
        a.step(10 * UP)
        b.step(10 * UP)
        done += 10 * UP
        counts.append(_spheres(a, wa))
        if done % 2000 == 0:
            chk = a.check_lists()
            print(f"  step {done}: {counts[-1]} spheres, {a.count_active()} contacts, list check {chk}", flush=True)
            assert not any(chk["violations"].values()), chk
    print(f"  step {done}: {counts[-1]} spheres, {a.count_active()} contacts", flush=True)
    assert all(x >= y for x, y in zip(counts, counts[1:])), counts
    assert counts[-1] < counts[0], "the falling bed crosses the zone's bound"
    sa, sb = a.state(), b.state()
    assert np.array_equal(a.models(), b.models())
    for f in ("pos", "vel", "av"):
        assert _bits(sa[f], sb[f]) == 0, f"{f}: two runs differ"
    assert np.isfinite(sa["pos"]).all() and np.isfinite(sa["vel"]).all()
    b.close()
    # restart at this step (a multiple of UpIact: a rebuild step)
    lo, hi, rec = a.active()
    print(f"  {len(lo)} active records at the restart", flush=True)
    assert len(lo) >= MIN_CONTACTS
    w2 = dataclasses.replace(wa, pos=sa["pos"], vel=sa["vel"], ox=sa["ox"], oz=sa["oz"], av=sa["av"],
                             model_of=a.models(), t_current=a.time)
    c = GpuSimulation(w2)
    assert c.n_do == a.n_do   # nothing more below the bound at the restart
    c.set_history(lo, hi, rec)
    a.step(3 * UP + 7)
    c.step(3 * UP + 7)
    s1, s2 = a.state(), c.state()
    assert np.array_equal(a.models(), c.models())
    for f in ("pos", "vel", "av"):
        assert _bits(s1[f], s2[f]) == 0, f"{f}: restart differs"
    print(f"  restart: bitwise equal after {3 * UP + 7} more steps ({a.n_do} DOs)", flush=True)
    for sim in (a, c):
        chk = sim.check_lists()
        assert not any(chk["violations"].values()), chk
    c.close()
    # the configuration's mixed precision: one FP32/FP64-mixed step against one
    # FP64 step from this state (contact histories carried), at full size, with
    # the bars of test_gpu_mixed.py
    from vedo_amd.precision import compare_precisions
    s3 = a.state()
    h3 = a.active()
    w3 = dataclasses.replace(wa, pos=s3["pos"], vel=s3["vel"], ox=s3["ox"], oz=s3["oz"], av=s3["av"],
                             model_of=a.models(), t_current=a.time)
    a.close()
    r = compare_precisions(w3, steps=1, history=h3)
    print(f"  mixed vs FP64, one step at {N} spheres: {r}", flush=True)
    assert r["active_pairs"] >= MIN_CONTACTS
    assert r["active_sets_identical"] and r["status_flags_identical"], r
    assert r["contact_err_max"] <= 1e-5, r
    assert r["sphere_impulse_err_max"] <= 1e-5 and r["sphere_ang_impulse_err_max"] <= 1e-5, r
