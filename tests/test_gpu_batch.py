"""Batched scenes (vg_batch_create / vg_admit): each scene's result equals the
same scene run alone (bitwise: a scene's grid, candidate order, walls, laws,
time step and rebuild cadence are its own), whether it is admitted at the
start or at a later rebuild boundary, and with other scenes finishing and
leaving the device around it.  The single-scene path is the reference's
(tests/test_gpu_parity.py), so this pins the batch to the reference too."""
import numpy as np
import pytest

from vedo_amd import scenes

pytestmark = pytest.mark.gpu

UP = 7


def _box(n, seed, law, **kw):
    args = dict(dt=2e-5, steps=400, spacing=2.02, jitter=0.05, vmax=0.2, offset=3.0, omega_max=5.0)
    args.update(dict(kn=2e3, e=0.6, mu=0.4) if law != "ISwH" else dict(E=1e6, nu=0.3, mu=0.5, e=0.5))
    args.update(kw)
    return scenes.box_scene(n, seed=seed, law=law, **args).to_arrays(up_iact=UP, record=1000)


def _alone(wa, steps):
    from vedo_amd.gpu import GpuSimulation
    sim = GpuSimulation(wa)
    sim.step(steps)
    st = sim.state()
    sim.close()
    return st


def _same(a, b, fields=("pos", "vel", "ox", "oz", "av")):
    return all(np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)) for f in fields)


def test_batch_scene_equals_scene_alone():
    from vedo_amd.gpu import SCENE_DONE, GpuBatch
    cases = [(_box(64, 11, "ISwBSDBF"), 40), (_box(27, 12, "ISwH", dt=1e-5), 33),
             (_box(125, 13, "ISwBSDBF", kn=5e3), 50), (_box(8, 14, "ISwBtSDBF"), 21)]
    b = GpuBatch(max_spheres=1000, max_walls=64, max_scenes=8, up_iact=UP)
    ids = [b.admit(wa, steps) for wa, steps in cases]
    print(f"\n  admitted {ids}", flush=True)
    late = _box(64, 15, "ISwBSDB2F")
    b.step(UP)                       # one boundary passes
    sid_late = b.admit(late, 30)     # placed at the next boundary
    for _ in range(12):
        b.step(UP)
    for (wa, steps), sid in zip(cases + [(late, 30)], ids + [sid_late]):
        state, done, _ = b.status(sid)
        assert state == SCENE_DONE and done == steps, (sid, state, done)
        got = b.state(sid)
        ref = _alone(wa, steps)
        sph = wa.model_shape[wa.model_of] == 1
        ok = _same({k: v[sph] for k, v in got.items()}, {k: v[sph] for k, v in ref.items()})
        print(f"  scene {sid}: {wa.n_do} DOs, {steps} steps, bitwise equal alone and batched: {ok}", flush=True)
        assert ok
        assert np.array_equal(got["pos"][~sph], ref["pos"][~sph])
    info = b.info()
    print(f"  batch after all finished: {info}", flush=True)
    assert info["on_device"] == 0 and info["spheres"] == 0
    b.close()


def test_batch_energy_and_active_match_scene_alone():
    from vedo_amd.gpu import GpuBatch, GpuSimulation
    wa = _box(64, 21, "ISwBSDBF")
    other = _box(125, 22, "ISwH", dt=1e-5)
    b = GpuBatch(max_spheres=1000, max_walls=64, max_scenes=4, up_iact=UP)
    s0 = b.admit(wa, -1)
    b.admit(other, -1)
    b.step(25)
    sim = GpuSimulation(wa)
    sim.step(25)
    eb, es = b.energy(s0), sim.system_energy()
    for k in ("potential", "translation", "rotation", "velocity_max", "mobile"):
        assert eb[k] == es[k], k
    lo, hi, _ = b.active(s0)
    lo2, hi2, _ = sim.active()
    assert sorted(zip(lo.tolist(), hi.tolist())) == sorted(zip(lo2.tolist(), hi2.tolist()))
    sim.close()
    b.close()


def test_batch_rejects_mismatched_up_iact_and_capacity():
    from vedo_amd.gpu import GpuBatch, GpuError
    b = GpuBatch(max_spheres=100, max_walls=16, max_scenes=2, up_iact=UP)
    wa = _box(64, 31, "ISwBSDBF")
    wa.up_iact = UP + 1
    with pytest.raises(GpuError, match="UpIact"):
        b.admit(wa, 10)
    wa.up_iact = UP
    b.admit(wa, 10)
    with pytest.raises(GpuError, match="capacity"):
        b.admit(_box(64, 32, "ISwBSDBF"), 10)
    b.close()


def test_batch_with_constrained_piston_equals_alone():
    # a constrained piston scene (its wall moves under its contacts' impulses,
    # summed in pair order) among scenes with fixed walls
    from conftest import load_golden
    from vedo_amd.gpu import SCENE_DONE, GpuBatch
    _, piston = load_golden("box_piston")
    cases = [(_box(64, 21, "ISwBSDBF"), 35), (piston, 42), (_box(27, 22, "ISwH", dt=1e-5), 28)]
    b = GpuBatch(max_spheres=1000, max_walls=64, max_scenes=4, up_iact=UP)
    ids = [b.admit(wa, steps) for wa, steps in cases]
    for _ in range(8):
        b.step(UP)
    for (wa, steps), sid in zip(cases, ids):
        state, done, _ = b.status(sid)
        assert state == SCENE_DONE and done == steps, (sid, state, done)
        got, ref = b.state(sid), _alone(wa, steps)
        ok = _same(got, ref, fields=("pos", "vel"))
        print(f"\n  scene {sid}: {wa.n_do} DOs (walls included), bitwise equal alone and batched: {ok}", flush=True)
        assert ok


def test_batch_with_zone_of_interest_equals_alone():
    # a scene losing spheres to its zone of interest (DO ids close up,
    # records re-keyed) among scenes that keep theirs
    from conftest import load_golden
    from vedo_amd.gpu import SCENE_DONE, GpuBatch
    _, zoi = load_golden("box_zoi")
    cases = [(_with_up(_box(64, 31, "ISwBSDBF"), 3), 24), (zoi, 30), (_with_up(_box(27, 32, "ISwBSDBF"), 3), 21)]
    b = GpuBatch(max_spheres=1000, max_walls=64, max_scenes=4, up_iact=3)
    ids = [b.admit(wa, steps) for wa, steps in cases]
    for _ in range(14):
        b.step(3)
    for (wa, steps), sid in zip(cases, ids):
        state, done, _ = b.status(sid)
        assert state == SCENE_DONE and done == steps, (sid, state, done)
        got, ref = b.state(sid), _alone(wa, steps)
        assert len(got["pos"]) == len(ref["pos"])
        ok = _same(got, ref, fields=("pos", "vel"))
        print(f"\n  scene {sid}: {len(got['pos'])} DOs left, bitwise equal alone and batched: {ok}", flush=True)
        assert ok


def _with_up(wa, up):
    import dataclasses
    return dataclasses.replace(wa, up_iact=up)


def test_graph_replayed_periods_equal_launched_steps():
    """Rebuild periods replayed from captured HIP graphs give bitwise the
    state of the same periods launched step by step (graphs off), period
    after period with the eager rebuilds between the replays, and with scenes
    finishing inside a period (launched eagerly) and at its end."""
    from vedo_amd.gpu import GpuBatch
    cases = [(_box(64, 31, "ISwBSDBF"), 7 * 9), (_box(125, 32, "ISwH", dt=1e-5), 7 * 6 + 3),
             (_box(27, 33, "ISwBtSDBF"), -1)]
    runs = []
    for graphs in (True, False):
        b = GpuBatch(max_spheres=1000, max_walls=64, max_scenes=4, up_iact=UP)
        b.set_graphs(graphs)
        ids = [b.admit(wa, steps) for wa, steps in cases]
        states = []
        for period in range(10):
            b.step(UP)
            states.append([b.state(s) if b.status(s)[0] != 0 else None for s in ids[2:]])
        chk = b.check_lists()
        assert not any(chk["violations"].values()), chk
        runs.append((b.graph_replays(), states, [b.state(s) for s in ids[:2]]))
        b.close()
    (rep_on, st_on, fin_on), (rep_off, st_off, fin_off) = runs
    print(f"\n  periods replayed from graphs: {rep_on} (graphs off: {rep_off})", flush=True)
    assert rep_on >= 5 and rep_off == 0
    for p, (x, y) in enumerate(zip(st_on, st_off)):
        assert _same(x[0], y[0]), f"running scene differs after period {p}"
    for x, y in zip(fin_on, fin_off):
        assert _same(x, y)
