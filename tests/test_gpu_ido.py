"""Restart on the GPU from a checkpoint the compiled reference wrote
(DOWorld::ReadIDO, DOWorld_RWIDO.cpp:9-55, read by vedo_amd.ido into
vg_create + vg_set_history), against the reference's own restart from that
file (tests/golden/ido_restart.npz, and ido_restart_pbc.npz: a box periodic
along x and y) at every dump step:
  * active contact set: bit-exact; flags, Kn, initial velocity, shear force
    of the active records: bit-exact (ISwBSDB2F, a spring-dashpot law);
  * positions, velocities, angular velocities: bit-exact; orientations to
    1e-12 relative (device cos/sin); last impulses to 1e-5 (the same bars
    as the fixture parity tests).
The .ido the GPU scene records after `record` restart steps equals the
reference's in every header, model, IactModel, name and record key byte;
its vectors meet the bars above, its record geometry (impact point,
direction, impulses to the master, overlap) 1e-12 relative."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _impulse_err(x, ref, tau=1e-3):
    nr = np.linalg.norm(ref, axis=1)
    scale = np.maximum(nr, tau * max(nr.max(), 1e-300))
    return float((np.linalg.norm(x - ref, axis=1) / scale).max())


def _rel(x, ref):
    return float(np.abs(np.asarray(x) - ref).max() / max(np.abs(ref).max(), 1e-300))


def _bits(a, b):
    return int((np.asarray(a).view(np.uint64) != np.asarray(b).view(np.uint64)).sum())


@pytest.mark.parametrize("name", ["ido_restart", "ido_restart_pbc"])
def test_gpu_restart_from_reference_ido(tmp_path, name):
    from vedo_amd.gpu import GpuSimulation
    from vedo_amd.ido import ido_history, ido_world, read_ido, record_ido
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    p0 = tmp_path / "start.ido"
    p0.write_bytes(z["ido0"].tobytes())
    p1 = tmp_path / "ref1.ido"
    p1.write_bytes(z["ido1"].tobytes())
    d0, d1 = read_ido(str(p0)), read_ido(str(p1))
    up, rec = (int(x) for x in z["cadence"])
    wa = ido_world(d0, up, rec)
    sim = GpuSimulation(wa)
    sim.set_history(*ido_history(d0))
    step = 0
    for k, dstep in enumerate(z["dumps"]):
        while step < dstep:
            sim.step(1)
            step += 1
            if step == rec:
                lo, hi, st = sim.active()
                mine = record_ido(str(tmp_path / "mine.ido"), wa, sim.state(), sim.time, lo, hi, st, template=d0)
                assert (mine.release, mine.max_id, mine.safety, mine.t_current, mine.dt) == \
                       (d1.release, d1.max_id, d1.safety, d1.t_current, d1.dt)
                assert [s.name for s in mine.statuses] == [s.name for s in d1.statuses]
                v, rv = np.stack([s.vectors for s in mine.statuses]), np.stack([s.vectors for s in d1.statuses])
                assert _bits(v[:, [0, 1, 4]], rv[:, [0, 1, 4]]) == 0
                assert _rel(v[:, 2:4], rv[:, 2:4]) <= 1e-12
                assert _impulse_err(v[:, 5], rv[:, 5]) <= 1e-5 and _impulse_err(v[:, 6], rv[:, 6]) <= 1e-5
                a, b = mine.records, d1.records
                for f in ("first", "second", "contact", "bond"):
                    assert np.array_equal(a[f], b[f]), f
                for f in ("kn", "init_vel", "shear"):
                    assert _bits(a[f], b[f]) == 0, f
                geo = {f: _rel(a[f], b[f]) for f in ("point", "dir", "to_master", "ang_to_master", "overlap")}
                print(f"\n  recorded .ido at restart step {step}: {len(a)} records, geometry rel {geo}", flush=True)
                assert max(geo.values()) <= 1e-12
                assert (tmp_path / "mine.ido").stat().st_size == len(z["ido1"])
        s = sim.state()
        ref = z[f"d{k}_state"]
        nb = {f: _bits(s[f], ref[:, i]) for i, f in ((0, "pos"), (1, "vel"), (4, "av"))}
        e_o = max(_rel(s["ox"], ref[:, 2]), _rel(s["oz"], ref[:, 3]))
        e_i = max(_impulse_err(s["impact"], ref[:, 5]), _impulse_err(s["ang_impact"], ref[:, 6]))
        lo, hi, st = sim.active()
        ri = z[f"d{k}_iact"]
        act = ri["active"].astype(bool)
        print(f"  restart step {dstep}: bitdiff {nb} orient {e_o:.1e} impulse {e_i:.1e} active {len(lo)}",
              flush=True)
        assert all(v == 0 for v in nb.values()) and e_o <= 1e-12 and e_i <= 1e-5
        assert sim.time == float(z[f"d{k}_t"])
        # both lists in key order (the reference's dump is in interaction order)
        rlo, rhi = np.minimum(ri["slave"], ri["master"])[act], np.maximum(ri["slave"], ri["master"])[act]
        ro, go = np.lexsort((rhi, rlo)), np.lexsort((hi, lo))
        lo, hi, st = lo[go], hi[go], st[go]
        assert np.array_equal(lo, rlo[ro]) and np.array_equal(hi, rhi[ro])
        rs = z[f"d{k}_status"][act][ro]
        assert np.array_equal(st["contact"], rs["flags"][:, 0]) and np.array_equal(st["last_contact"], rs["flags"][:, 1])
        assert np.array_equal(st["bond"], rs["flags"][:, 2]) and np.array_equal(st["last_bond"], rs["flags"][:, 3])
        for f in ("kn", "init_vel", "shear"):
            assert _bits(st[f], rs[f]) == 0, f
    sim.close()
