"""Batched and served scenes against the reference's own golden vectors.

The golden fixtures (tests/golden, written by the compiled reference) with
UpIact 7 run together in one batch (vg_batch_create / vg_admit) and through
the continuous-batching server with a capacity below their total size, so
that later scenes wait for slots and are placed at later rebuild boundaries.
The plate with a circular hole of box_hole makes the whole batch run the
kernels that read the hole table, the other scenes' walls marked hole-free.
Each scene's state is compared directly with the reference's dumps:
bit-exact for the spring-dashpot laws, relative 1e-9 for ISwH / ISwHertz
(device pow), as in tests/test_gpu_parity.py.  A full-size property test
runs the batched configuration's 256 box scenes (10.8M spheres) for 200
steps and checks each scene bitwise against the same scene alone."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

UP = 7
CASES = ["box_bsdbf", "box_h", "box_hertz", "box_btsdbf", "box_b2f", "box_bond", "box_bond_bt", "box_cyl",
         "box_orbital", "box_piston", "box_hole", "box_crowd"]
EXACT_LAWS = {1, 2, 3}


def _check(name, z, wa, st, k):
    ref = z[f"d{k}_state"]
    sph = wa.model_shape[wa.model_of] == 1
    laws = set(int(x) for x in np.unique(wa.iact_law) if x)
    if laws <= EXACT_LAWS:
        for f, i in (("pos", 0), ("vel", 1), ("av", 4)):
            nb = int((st[f][sph] != ref[sph, i]).sum())
            assert nb == 0, f"{name} dump {int(z['dumps'][k])}: {f} differs in {nb} values"
    else:
        for f, i in (("pos", 0), ("vel", 1), ("av", 4)):
            e = np.abs(st[f][sph] - ref[sph, i]).max() / max(np.abs(ref[sph, i]).max(), 1e-300)
            assert e <= 1e-9, (name, int(z["dumps"][k]), f, e)


def test_batch_of_fixture_scenes_matches_reference_dumps():
    from vedo_amd.gpu import GpuBatch
    fx = [(name,) + load_golden(name) for name in CASES]
    for _, z, wa in fx:
        assert int(wa.up_iact) == UP
    b = GpuBatch(max_spheres=2000, max_walls=256, max_scenes=16, up_iact=UP)
    ids = [b.admit(wa, int(z["dumps"][-1])) for _, z, wa in fx]
    checked = 0
    last = max(int(z["dumps"][-1]) for _, z, _ in fx)
    for step in range(1, last + 1):
        b.step(1)
        for (name, z, wa), sid in zip(fx, ids):
            hit = np.nonzero(np.asarray(z["dumps"]) == step)[0]
            if len(hit):
                _check(name, z, wa, b.state(sid), int(hit[0]))
                checked += 1
    print(f"\n  {len(fx)} fixture scenes in one batch: {checked} dumps equal the reference's", flush=True)
    assert checked == sum(len(z["dumps"]) - 1 for _, z, _ in fx)
    b.close()


def test_server_results_match_reference_final_dumps():
    from vedo_amd import serving
    fx = [(name,) + load_golden(name) for name in CASES]
    reqs = []
    for k, (name, z, wa) in enumerate(fx):
        n = int((wa.model_shape[wa.model_of] == 1).sum())
        reqs.append(serving.Request(k, 0.01 * k, "box", n, int(z["dumps"][-1]), "fixture", 0, wa=wa))
    # room for three scenes at a time: admission by footprint binds
    cap = 3 * max(r.n for r in reqs)
    srv = serving.Server(capacity=cap, up_iact=UP, max_walls=256, max_scenes=16)
    st = srv.run(reqs)
    srv.close()
    # placed only after an earlier scene had finished and left its slots
    first_done = min(r.t_done for r in reqs)
    waited = sum(1 for r in reqs if r.t_admit >= first_done)
    print(f"\n  served {len(reqs)} fixture scenes with room for 3 at a time ({waited} waited for slots), "
          f"{st.failed} failed", flush=True)
    assert st.failed == 0 and waited > 0
    for r, (name, z, wa) in zip(reqs, fx):
        _check(name, z, wa, r.result, len(z["dumps"]) - 1)


def test_batched_configuration_scenes_equal_alone():
    """256 scenes, 10.8M spheres, 200 steps (10 rebuild periods): every scene
    of the batch bitwise equal to the same scene run alone, none failed."""
    from vedo_amd import scenes
    from vedo_amd.gpu import SCENE_DONE, GpuBatch, GpuSimulation
    steps, up = 200, 20
    was = [scenes.batched_box(k, steps, up) for k in range(256)]
    n_of = [int((wa.model_shape[wa.model_of] == 1).sum()) for wa in was]
    b = GpuBatch(sum(n_of), max_walls=6 * 256, max_scenes=256, up_iact=up)
    ids = [b.admit(wa, steps) for wa in was]
    for _ in range(steps // up + 1):
        b.step(up)
        chk = b.check_lists()
        assert not any(chk["violations"].values()), chk
    got = []
    for sid in ids:
        state, done, _ = b.status(sid)
        assert state == SCENE_DONE and done == steps, (sid, state, done)
        got.append(b.state(sid))
    b.close()
    print(f"\n  batch of {len(was)} scenes ({sum(n_of)} spheres) done", flush=True)
    for k, wa in enumerate(was):
        sim = GpuSimulation(wa)
        sim.step(steps)
        ref = sim.state()
        sim.close()
        for f in ("pos", "vel", "av"):
            assert np.array_equal(got[k][f].view(np.uint64), ref[f].view(np.uint64)), (k, n_of[k], f)
        if k % 64 == 63:
            print(f"  scenes 0..{k} bitwise equal alone", flush=True)
