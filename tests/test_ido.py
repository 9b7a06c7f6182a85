"""The reference's .ido checkpoint (DOWorld::ReadIDO / WriteIDO,
DOWorld_RWIDO.cpp:9-202) in vedo_amd.ido, against two files the compiled
reference wrote (tests/golden/make_golden.py, make_ido_restart): the reader
and writer round-trip both byte for byte; a scene read from the first with
its IactRecordTab continues in the oracle bit for bit as the reference's own
restart from that file does (DOWorld::ReadIDO, then SimMediator::Run); and
the .ido the oracle's restarted scene records after `record` steps equals,
byte for byte, the one the reference's restarted run wrote.  The same for
a periodic box (fixture ido_restart_pbc), whose files carry the periodic
boundary."""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN

FIELDS = ("pos", "vel", "ox", "oz", "av", "impact", "ang_impact")


RESTARTS = ["ido_restart", "ido_restart_pbc"]
N_DO = {"ido_restart": 70, "ido_restart_pbc": 146}


def _z(name="ido_restart"):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def _scene(name):
    """The scene each restart fixture started from (make_ido_restart)."""
    from vedo_amd import scenes
    if name == "ido_restart_pbc":
        return scenes.periodic_box((6, 6, 4), "xy", seed=34, law="ISwBSDBF", vmax=0.6)
    return scenes.box_scene(64, seed=18, law="ISwBSDB2F", dt=2e-5, steps=400, omega_max=5.0, spacing=1.99,
                            jitter=0.01, vmax=0.02, offset=0.0, kn=2e3, e=0.6, mu=0.4)


def _file(tmp_path, z, key):
    p = tmp_path / f"{key}.ido"
    p.write_bytes(z[key].tobytes())
    return str(p)


@pytest.mark.parametrize("name", RESTARTS)
@pytest.mark.parametrize("key", ["ido0", "ido1"])
def test_ido_round_trip_is_byte_identical(tmp_path, key, name):
    from vedo_amd.ido import read_ido, write_ido
    z = _z(name)
    d = read_ido(_file(tmp_path, z, key))
    out = tmp_path / "out.ido"
    write_ido(str(out), d)
    assert out.read_bytes() == z[key].tobytes()
    cad = z["cadence"]
    assert d.release == "2016-09-23" and d.udd_dos == 0 and d.udd_is == 0
    assert len(d.statuses) == N_DO[name] and len(d.records) > 0
    if name == "ido_restart_pbc":
        # periodic along x and y (SystemParameter::operator>>, SystemParameter.cpp:87-89)
        assert bytes(d.pbc_switch) == b"\x01\x01\x00" and d.pbc_hi[2] == 0.0 and d.pbc_hi[0] > 0.0
    # records are in the std::map's key order: (first, second) ascending
    keys = d.records["first"].astype(np.int64) << 32 | d.records["second"]
    assert np.all(np.diff(keys) > 0)
    assert d.t_current == pytest.approx(int(cad[1]) * d.dt * (1 if key == "ido0" else 2), rel=1e-12)


def test_ido_reader_rejects_bad_files(tmp_path):
    from vedo_amd.ido import read_ido
    b = _z()["ido0"].tobytes()
    p = tmp_path / "bad.ido"
    p.write_bytes(b[:-7])
    with pytest.raises(ValueError, match="truncated"):
        read_ido(str(p))
    p.write_bytes(b + b"\0")
    with pytest.raises(ValueError, match="trailing"):
        read_ido(str(p))
    # the integer-count field after the release string (SystemParameter.cpp:101-108)
    n = struct.unpack_from("<I", b, 0)[0]
    p.write_bytes(b[:4 + n] + struct.pack("<I", 4) + b[8 + n:])
    with pytest.raises(ValueError, match="integer"):
        read_ido(str(p))


def test_ido_history_mirrors_the_reader():
    """IactRecordTab::operator<< (IactRecordTab.cpp:453-504) fills one
    ImpactStatus whose constructor sets LastContact and LastBond
    (ImpactStatus.cpp:8-9) and assigns only Contact and Bond."""
    import tempfile
    from vedo_amd.ido import ido_history, read_ido
    z = _z()
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "a.ido")
        with open(p, "wb") as f:
            f.write(z["ido0"].tobytes())
        d = read_ido(p)
    lo, hi, st = ido_history(d)
    assert (st["last_contact"] == 1).all() and (st["last_bond"] == 1).all()
    assert np.array_equal(st["contact"], d.records["contact"]) and np.array_equal(lo, d.records["first"])
    assert st["contact"].sum() > 50


def restart_and_compare(make_sim, tmp_path, name="ido_restart"):
    """Run the restart of `ido0` on make_sim(WorldArrays) and compare with the
    reference's dumps and its first recorded .ido."""
    from vedo_amd.ido import ido_history, ido_world, read_ido, record_ido
    z = _z(name)
    d0 = read_ido(_file(tmp_path, z, "ido0"))
    up, rec = (int(x) for x in z["cadence"])
    wa = ido_world(d0, up, rec)
    sim = make_sim(wa)
    sim.set_history(*ido_history(d0))
    step = 0
    for k, dstep in enumerate(z["dumps"]):
        while step < dstep:
            sim.step(1)
            step += 1
            if step == rec:
                lo, hi, st = sim.active()
                out = tmp_path / "mine.ido"
                record_ido(str(out), wa, sim.state(), sim.time, lo, hi, st, template=d0)
                assert out.read_bytes() == z["ido1"].tobytes(), "recorded .ido differs"
        s = sim.state()
        mine = np.stack([s[f] for f in FIELDS], 1)
        ref = z[f"d{k}_state"]
        assert np.array_equal(mine.view(np.uint64), ref.view(np.uint64)), f"state at restart step {dstep}"
        assert sim.time == float(z[f"d{k}_t"])
        lo, hi, st = sim.active()
        ri = z[f"d{k}_iact"]
        act = ri["active"].astype(bool)
        np.testing.assert_array_equal(lo, np.minimum(ri["slave"], ri["master"])[act])
        np.testing.assert_array_equal(hi, np.maximum(ri["slave"], ri["master"])[act])
        rs = z[f"d{k}_status"][act]
        for f in ("kn", "init_vel", "shear", "point", "dir", "to_master", "ang_to_master", "overlap"):
            assert np.array_equal(np.asarray(st[f]).view(np.uint64), np.asarray(rs[f]).view(np.uint64)), f
        np.testing.assert_array_equal(st["contact"], rs["flags"][:, 0])
        np.testing.assert_array_equal(st["last_contact"], rs["flags"][:, 1])
        np.testing.assert_array_equal(st["bond"], rs["flags"][:, 2])
        np.testing.assert_array_equal(st["last_bond"], rs["flags"][:, 3])
    return sim


@pytest.mark.parametrize("name", RESTARTS)
def test_oracle_restart_from_reference_ido(tmp_path, name):
    from oracle.oracle import OracleSim
    restart_and_compare(OracleSim, tmp_path, name)


@pytest.mark.parametrize("name", RESTARTS)
def test_record_ido_from_the_scene_matches_reference(tmp_path, name):
    """The .ido the oracle records after 10 steps of the fixture's scene
    (rebuilt here with the seeded generator that made it), with models and
    IactModels taken from the scene description, equals byte for byte the
    first file the reference recorded (Consultant::RecordIDO)."""
    from oracle.oracle import OracleSim
    from vedo_amd.ido import record_ido
    z = _z(name)
    up, rec = (int(x) for x in z["cadence"])
    wa = _scene(name).to_arrays(up_iact=up, record=rec)
    o = OracleSim(wa)
    o.step(rec)
    out = tmp_path / "mine.ido"
    record_ido(str(out), wa, o.state(), o.time, *o.active())
    assert out.read_bytes() == z["ido0"].tobytes()


def test_record_ido_after_zone_of_interest_erasures(tmp_path):
    """After Consultant::CleanUp erased spheres (box_zoi), the remaining DOs
    are renumbered: the recorded .ido names each by its own model (the walls
    stay walls), reads back with the same shapes, and a record without the
    per-DO models is refused instead of naming walls after the sphere model."""
    from conftest import load_golden
    from oracle.oracle import OracleSim
    from vedo_amd.ido import ido_world, read_ido, record_ido
    from vedo_amd.world import VedoError
    z, wa = load_golden("box_zoi")
    # the fixture's scene: the sphere model and the box's plate model
    assert list(wa.model_shape) == [1, 7]
    wa.model_names, wa.group_names = ["Sphere", "Plate"], ["Granule", "Wall"]
    o = OracleSim(wa)
    o.step(4)
    ids = o.ids()
    assert len(ids) < wa.n_do, "the zone of interest erased nothing"
    model_of = wa.model_of[ids]
    with pytest.raises(VedoError):
        record_ido(str(tmp_path / "bad.ido"), wa, o.state(), o.time, *o.active())
    out = tmp_path / "zoi.ido"
    record_ido(str(out), wa, o.state(), o.time, *o.active(), model_of=model_of)
    d = read_ido(str(out))
    names = wa.model_names
    assert [s.name for s in d.statuses] == [names[int(m)] for m in model_of]
    back = ido_world(d, int(wa.up_iact), int(wa.record))
    assert back.n_do == len(ids)
    assert np.array_equal(back.model_shape[back.model_of], wa.model_shape[model_of])
    assert int((back.model_shape[back.model_of] != 1).sum()) == int((wa.model_shape[wa.model_of] != 1).sum())
    np.testing.assert_array_equal(back.pos, o.state()["pos"])
