"""Generate golden vectors from the compiled reference (oracle/_ref).

Run in the build container, where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py

For every case in CASES it writes the scene exchange file, runs
oracle/_ref/vedo_ref_harness (the reference's own SimMediator/Consultant/
ImpactSolver code) in a scratch directory, and stores the dumps as
tests/golden/<case>.npz: inputs (the scene arrays) and expected outputs
(per-DO state and impulses, interaction list, contact depths, ImpactStatus of
every interaction) at the listed steps.  The fixtures are data only.
"""
from __future__ import annotations

import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vedo_amd import scenes  # noqa: E402
from vedo_amd.world import write_vscene  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "vedo_ref_harness")

STATUS_FIELDS = ("kn", "init_vel", "shear", "point", "dir", "to_master", "ang_to_master", "overlap")


def parse_dumps(path):
    out = []
    with open(path, "rb") as f:
        buf = f.read()
    o = 0
    while o < len(buf):
        assert buf[o:o + 4] == b"DUMP", "bad dump magic"
        o += 4
        step, = struct.unpack_from("<i", buf, o); o += 4
        t, = struct.unpack_from("<d", buf, o); o += 8
        n, = struct.unpack_from("<i", buf, o); o += 4
        st = np.frombuffer(buf, "<f8", n * 21, o).reshape(n, 7, 3).copy(); o += n * 21 * 8
        ni, = struct.unpack_from("<i", buf, o); o += 4
        rec = np.frombuffer(buf, np.dtype([("slave", "<u4"), ("master", "<u4"), ("active", "<u4"),
                                           ("depth", "<f8")]), ni, o).copy(); o += ni * 20
        sdt = np.dtype([("flags", "u1", 4), ("kn", "<f8"), ("init_vel", "<f8"), ("shear", "<f8", 3),
                        ("point", "<f8", 3), ("dir", "<f8", 3), ("to_master", "<f8", 3),
                        ("ang_to_master", "<f8", 3), ("overlap", "<f8")])
        ist = np.frombuffer(buf, sdt, ni, o).copy(); o += ni * sdt.itemsize
        out.append(dict(step=step, t=t, state=st, iact=rec, status=ist))
    return out


def _wall_box_ids(wa):
    return np.nonzero(wa.model_shape[wa.model_of] != 1)[0]


def make_cases():
    """(name, WorldArrays, consultant, nsteps, dumps, assigned-statuses or None)."""
    cases = []
    common = dict(spacing=2.02, jitter=0.05, vmax=0.2, offset=3.0)
    for law, name, extra in (("ISwBSDBF", "box_bsdbf", dict(kn=2e3, e=0.6, mu=0.4)),
                             ("ISwH", "box_h", dict(E=1e6, nu=0.3, mu=0.5, e=0.5)),
                             ("ISwHertz", "box_hertz", dict(E=1e6, nu=0.3)),
                             ("ISwBtSDBF", "box_btsdbf", dict(kn=2e3, e=0.6, mu=0.4))):
        w = scenes.box_scene(64, seed=11, law=law, dt=2e-5, steps=400, omega_max=5.0, **common, **extra)
        wa = w.to_arrays(up_iact=7, record=1000)
        cases.append((name, wa, 0, 150, [0, 1, 2, 6, 7, 8, 50, 150], None))
    # the spring-dashpot laws with normal and shear bonds: a lattice a hair
    # apart, so that separated neighbours lie within the bond reach bn/kn,
    # with jitter (some pairs overlap) and random velocities that pull bonds
    # into tension until they break; ISwBSDB2F applies the bonded tension
    # only on its LastBond branch (fresh interactions at the start and after
    # every rebuild), so the three laws give three different trajectories
    for law, name in (("ISwBSDB2F", "box_b2f"), ("ISwBSDBF", "box_bond"), ("ISwBtSDBF", "box_bond_bt")):
        cases.append((name, bond_world(law).to_arrays(up_iact=7, record=1000), 0, 150,
                      [0, 1, 2, 6, 7, 8, 50, 150], None))
    # quasi-cylinders: a fixed rod under the bed, shorter than the bed (the
    # spheres beyond its ends meet the end caps), and an orbital rod that
    # rises into the bed while it turns (DOOrbitalQuasiCylinder)
    cases.append(("box_cyl", cyl_world().to_arrays(up_iact=7, record=1000), 0, 200,
                  [0, 1, 2, 6, 7, 8, 50, 100, 150, 200], None))
    # orbital quasi-plates (DOOrbitalQuasiPlate) that move by themselves: a
    # floor rising and turning about a tilted axis, a side wall pushing in
    cases.append(("box_orbital", orbital_world().to_arrays(up_iact=7, record=1000), 0, 150,
                  [0, 1, 2, 6, 7, 8, 50, 100, 150], None))
    # crowded lists: two large spheres each touched by about 100 small ones
    # (more than 64 entries in one sphere's neighbour list)
    cases.append(("box_crowd", crowd_world().to_arrays(up_iact=7, record=1000), 0, 150,
                  [0, 1, 2, 6, 7, 8, 50, 100, 150], None))
    # NBS consultant on the same kind of scene
    w = scenes.box_scene(64, seed=12, law="ISwBSDBF", dt=2e-5, steps=400, kn=2e3, e=0.6, mu=0.4, **common)
    cases.append(("box_nbs", w.to_arrays(up_iact=5, record=1000), 1, 100, [0, 1, 5, 6, 100], None))
    # NBS's quirks: every coordinate negative (its maxima start from
    # numeric_limits<double>::min()) and a smaller last sphere (its zone size
    # takes the last local DO's range)
    cases.append(("box_nbs_quirks", nbs_quirk_world().to_arrays(up_iact=5, record=1000), 1, 100,
                  [0, 1, 5, 6, 10, 11, 50, 100], None))
    # polydisperse drum section with kinematic walls (prescribed rigid rotation)
    w, wall_ids, c, a = scenes.drum_scene(300, seed=13, R=0.02, L=0.012, omega=20.0, dt=1e-5, steps=500,
                                          sides=16, rmin=0.6e-3, rmax=1.4e-3, classes=5, kn=2e3)
    wa = w.to_arrays(up_iact=9, record=1000)
    nsteps = 120
    ids = np.array(wall_ids, np.uint32)
    p0, v0, ox0, oz0 = wa.pos[ids], wa.vel[ids], wa.ox[ids], wa.oz[ids]
    assigned = []
    for k in range(1, nsteps + 1):
        p, v, ox, oz, av = scenes.rigid_rotation_status(p0, v0, ox0, oz0, c, a, 20.0, k * wa.dt)
        assigned.append(np.concatenate([p, v, ox, oz, av], axis=1))
    cases.append(("drum_kin", wa, 0, nsteps, [0, 1, 2, 9, 10, 60, 120], (ids, np.stack(assigned))))
    # Run(map) on spheres: a few spheres take new statuses (moved by up to
    # 0.2 r, new velocities) at some steps, one of them a rebuild step; rows
    # of NaN are "not assigned at this step"
    w = scenes.box_scene(64, seed=14, law="ISwBSDBF", dt=2e-5, steps=400, kn=2e3, e=0.6, mu=0.4,
                         omega_max=5.0, **common)
    wa = w.to_arrays(up_iact=7, record=1000)
    nsteps = 40
    ids = np.array([5, 17, 40, 41], np.uint32)
    rng = np.random.default_rng(14)
    r = wa.model_attr[wa.model_of[ids], 0]
    rows = np.full((nsteps, len(ids), 15), np.nan)
    for step in (3, 7, 8, 21):
        sel = rng.random(len(ids)) < 0.75
        sel[0] = True
        p = wa.pos[ids] + rng.uniform(-0.2, 0.2, (len(ids), 3)) * r[:, None]
        v = rng.uniform(-0.3, 0.3, (len(ids), 3))
        th = rng.uniform(0, 2 * np.pi, len(ids))
        ox = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1)
        oz = np.tile([0.0, 0.0, 1.0], (len(ids), 1))
        av = rng.uniform(-5, 5, (len(ids), 3))
        st = np.concatenate([p, v, ox, oz, av], 1)
        rows[step - 1][sel] = st[sel]
    cases.append(("box_assign", wa, 0, nsteps, [0, 2, 3, 4, 7, 8, 9, 21, 22, 40], (ids, rows)))
    # a constrained piston (DOConstrainedQuasiPlate): the top plate of a small
    # box rests on the top layer of spheres and presses down under its weight
    # and an external force, moving along its Oz
    w = scenes.box_scene(64, seed=15, law="ISwBSDBF", dt=2e-5, steps=400, kn=2e3, e=0.6, mu=0.4, side=0.0125,
                         spacing=2.02, jitter=0.05, vmax=0.05, offset=3.0, omega_max=5.0)
    cases.append(("box_piston", piston_world(w).to_arrays(up_iact=7, record=1000), 0, 150,
                  [0, 1, 2, 6, 7, 8, 50, 150], None))
    # three constrained walls meeting at a corner (the piston and the x-min
    # and y-min side plates, DOConstrainedQuasiPlate): the corner spheres have
    # all three among their candidates, and those at the top touch them all
    w = scenes.box_scene(64, seed=18, law="ISwBSDBF", dt=2e-5, steps=400, kn=2e3, e=0.6, mu=0.4, side=0.0125,
                         spacing=2.02, jitter=0.05, vmax=0.05, offset=3.0, omega_max=5.0)
    cases.append(("box_cw3", corner_world(piston_world(w)).to_arrays(up_iact=7, record=1000), 0, 150,
                  [0, 1, 2, 6, 7, 8, 50, 150], None))
    # a zone of interest: the box has no floor, the spheres move down and
    # the lowest leave the zone while touching the layer above; they are
    # erased at the next rebuilds (Consultant::CleanUp) and the records
    # move with IactRecordTab::ModifyPair
    # (pre-compressed lattice moving down as one: the lowest layer leaves the
    # zone at the first rebuild still pressed against the layer above, so
    # records of erased spheres move onto other pairs)
    w = scenes.box_scene(64, seed=16, law="ISwBSDBF", dt=2e-5, steps=400, kn=2e3, e=0.6, mu=0.4,
                         spacing=1.98, jitter=0.0, vmax=0.0, offset=3.0)
    cases.append(("box_zoi", zoi_world(w, vz=-0.4, below=0.01).to_arrays(up_iact=3, record=1000), 0, 60,
                  [0, 2, 3, 4, 6, 7, 30, 60], None))
    # a floor with a circular hole (CDSphere_QuasiPlateWithCircularHole)
    w = scenes.box_scene(64, seed=17, law="ISwBSDBF", dt=2e-5, steps=400, kn=2e3, e=0.6, mu=0.4, **common)
    cases.append(("box_hole", hole_world(w).to_arrays(up_iact=7, record=1000), 0, 200,
                  [0, 1, 2, 6, 7, 8, 50, 100, 200], None))
    # periodic boundaries (PeriodicBoundaryConditions): lattices that fill
    # the periodic cell, so the outer layers meet across the boundary and
    # spheres leave through one face and come back through the other.  Two
    # periodic axes with a floor and a lid (ISwBSDBF), three periodic axes
    # without walls or gravity (ISwBtSDBF), one periodic axis between walls
    # (ISwHertz), two with ISwH; the Leap traversal's wrapped half-shell
    # cells differ with the number of periodic axes.  The pow-based laws run
    # 150 and 60 steps (their device pow rounds differently, and a dense
    # ISwH bed amplifies the difference a hundredfold every 50 steps)
    cases.append(("box_pbc_xy", scenes.periodic_box((6, 6, 4), "xy", seed=31, law="ISwBSDBF", vmax=0.6)
                  .to_arrays(up_iact=7, record=1000), 0, 300, [0, 1, 2, 6, 7, 8, 50, 100, 200, 300], None))
    cases.append(("box_pbc_xyz", scenes.periodic_box((5, 5, 5), "xyz", seed=32, law="ISwBtSDBF", vmax=0.6)
                  .to_arrays(up_iact=7, record=1000), 0, 300, [0, 1, 2, 6, 7, 8, 50, 100, 200, 300], None))
    cases.append(("box_pbc_x", scenes.periodic_box((7, 4, 4), "x", seed=33, law="ISwHertz", E=1e6, vmax=1.0)
                  .to_arrays(up_iact=7, record=1000), 0, 150, [0, 1, 2, 6, 7, 8, 50, 100, 150], None))
    cases.append(("box_pbc_h", scenes.periodic_box((5, 5, 4), "xy", seed=36, law="ISwH", E=1e6, vmax=1.0,
                                                    drift=(1.5, 1.0, 0.0))
                  .to_arrays(up_iact=7, record=1000), 0, 60, [0, 1, 2, 6, 7, 8, 14, 15, 28, 29, 49, 50, 60], None))
    # NBSConsultant's candidate list in a periodic cell: its zones never wrap
    # (pairs across the boundary are not candidates), while contact detection
    # of the listed pairs and the positions do
    cases.append(("box_pbc_nbs", scenes.periodic_box((6, 6, 4), "xy", seed=35, law="ISwBSDBF", vmax=0.6,
                                                      drift=(1.0, 0.5, 0.0))
                  .to_arrays(up_iact=5, record=1000), 1, 200, [0, 1, 4, 5, 6, 50, 100, 200], None))
    return cases


def bond_world(law, bn=0.03, bs=2e-3, seed=21):
    """64 spheres 2.003 r apart (jitter 0.004 r, speeds up to 0.15 m/s,
    spins up to 5 rad/s) with NormalBondStrength bn (bond reach bn/kn =
    0.015 r at kn = 2e3) and ShearBondStrength bs on every IactModel."""
    w = scenes.box_scene(64, seed=seed, law=law, dt=2e-5, steps=400, omega_max=5.0, spacing=2.003,
                         jitter=0.004, vmax=0.15, offset=3.0, kn=2e3, e=0.6, mu=0.4)
    for im in w.iact_models:
        im.mechanisms["NormalBondStrength"] = bn
        im.mechanisms["ShearBondStrength"] = bs
    return w


def nbs_quirk_world(seed=24, shift=-0.05, small=0.7):
    """64 spheres in their box moved to negative coordinates, every fourth
    sphere (the last one among them) a smaller model of radius small*r."""
    from vedo_amd.world import DOModel, DOStatus, IactModel
    w = scenes.box_scene(64, seed=seed, law="ISwBSDBF", dt=2e-5, steps=400, omega_max=5.0, spacing=2.02,
                         jitter=0.05, vmax=0.3, offset=3.0, kn=2e3, e=0.6, mu=0.4)
    big = [m for m in w.models if m.name == "Sphere"][0]
    r = big.attributes[0]
    w.models.append(DOModel("Small", big.group, "mobile", "local", big.density, 1.0, (0, 0, 0), "Sphere",
                            (small * r,)))
    sph = [k for k, st in enumerate(w.statuses) if st.name == "Sphere"]
    for q, k in enumerate(sph):
        st = w.statuses[k]
        name = "Small" if (q % 4 == 3) else st.name
        w.statuses[k] = DOStatus(name, tuple(np.asarray(st.position) + shift), st.velocity, st.orientation_x,
                                 st.orientation_z, st.angular_velocity)
    for k, st in enumerate(w.statuses):
        if st.name == "Plate":
            w.statuses[k] = DOStatus(st.name, tuple(np.asarray(st.position) + shift), st.velocity,
                                     st.orientation_x, st.orientation_z, st.angular_velocity)
    return w


def cyl_world(seed=22, vz=-0.25):
    """A falling lattice over a fixed quasi-cylinder rod along x, shorter
    than the lattice so that the spheres past its ends touch the caps, and
    an orbital rod along y moving up at 0.15 m/s while it turns."""
    from vedo_amd.world import DOModel, DOStatus
    w = scenes.box_scene(64, seed=seed, law="ISwBSDBF", dt=2e-5, steps=400, omega_max=5.0, spacing=2.02,
                         jitter=0.05, vmax=0.1, offset=3.0, kn=2e3, e=0.6, mu=0.4)
    r = [m for m in w.models if m.name == "Sphere"][0].attributes[0]
    sph = [k for k, st in enumerate(w.statuses) if st.name == "Sphere"]
    p = np.array([w.statuses[k].position for k in sph])
    lo, hi = p.min(0), p.max(0)
    rc = 0.6 * r
    w.models.append(DOModel("Rod", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiCylinder",
                            (rc, 0.36 * (hi[0] - lo[0]))))
    w.statuses.append(DOStatus("Rod", (0.5 * (lo[0] + hi[0]) + 0.2 * r, 0.5 * (lo[1] + hi[1]) + 0.3 * r,
                                       lo[2] - r - rc - 0.05 * r), (0, 0, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0)))
    w.models.append(DOModel("Spinner", "Wall", "orbital", "global", 0.0, 1.0, (0, 0, 0), "QuasiCylinder",
                            (0.5 * r, 0.3 * (hi[1] - lo[1]))))
    w.statuses.append(DOStatus("Spinner", (lo[0] + r, 0.5 * (lo[1] + hi[1]) - 0.4 * r, lo[2] - 2.2 * r),
                               (0.0, 0.0, 0.15), (1, 0, 0), (0, 1, 0), (0.8, 6.0, 0.0)))
    for k in sph:
        st = w.statuses[k]
        w.statuses[k] = DOStatus(st.name, st.position, (st.velocity[0], st.velocity[1], st.velocity[2] + vz),
                                 st.orientation_x, st.orientation_z, st.angular_velocity)
    return w


def crowd_world(seed=24, n_shell=100, r=0.5e-3, big=3e-3, dt=5e-6, steps=400):
    """Two large spheres (radius 6 r) in a box, each wrapped in a shell of
    `n_shell` small spheres (Fibonacci points 0.01 r off its surface, at least
    2.16 r apart) moving onto it at 0.1 m/s, under gravity: each large sphere
    lists about 100 partners and is touched by all of them at once (more
    than 64 list entries and history slots)."""
    from vedo_amd.world import DOModel, DOStatus, DOWorld, IactModel, SystemParameter
    rng = scenes.SplitMix64(seed)
    d = big + r + 0.01 * r
    k = np.arange(n_shell) + 0.5
    z = 1.0 - 2.0 * k / n_shell
    phi = np.pi * (1.0 + 5.0 ** 0.5) * k
    u = np.stack([np.sqrt(1.0 - z * z) * np.cos(phi), np.sqrt(1.0 - z * z) * np.sin(phi), z], 1)
    gap = np.linalg.norm(u[:, None] - u[None], axis=2) * d + 4.0 * r * np.eye(n_shell)
    assert gap.min() > 2.0 * r, "the shell's spheres overlap"
    m = d + r + 2.0 * r   # a shell's reach from its centre, and a margin to the walls
    centres = [(m, m, m + 0.2 * r), (3.0 * m + 2.0 * r, m + 0.3 * r, m)]
    models = [DOModel("Big", "Granule", "mobile", "local", 2500.0, 1.0, (0, 0, 0), "Sphere", (big,)),
              DOModel("Sphere", "Granule", "mobile", "local", 2500.0, 1.0, (0, 0, 0), "Sphere", (r,))]
    statuses = [DOStatus("Big", c, (0.0, 0.0, 0.0), (1, 0, 0), (0, 0, 1), (0.0, 0.0, 0.0)) for c in centres]
    for c in centres:
        v = rng.uniform(3 * n_shell, -0.02, 0.02).reshape(n_shell, 3) - 0.1 * u
        w = rng.uniform(3 * n_shell, -5.0, 5.0).reshape(n_shell, 3)
        for q in range(n_shell):
            statuses.append(DOStatus("Sphere", tuple(np.array(c) + d * u[q]), tuple(v[q]), (1, 0, 0), (0, 0, 1),
                                     tuple(w[q])))
    wm, ws = scenes.box_walls(4.0 * m + 2.0 * r, 2e-3)
    iact = [IactModel("Granule", "Granule", "ISwBSDBF", scenes.spring_dashpot(2e3, 0.8, 0.6, 0.4, False)),
            IactModel("Wall", "Granule", "ISwBSDBF", scenes.spring_dashpot(2e3, 0.8, 0.6, 0.4, False))]
    sp = SystemParameter(0.0, steps * dt, dt, 0.0, (0.0, 0.0, -9.81))
    return DOWorld(sp, models + wm, iact, statuses + ws)


def orbital_world(seed=23):
    """The box's floor as an orbital plate rising at 0.3 m/s and turning at
    (0.4, -0.3, 5) rad/s, and its x = 0 side as one moving in at 0.2 m/s."""
    from vedo_amd.world import DOModel, DOStatus
    w = scenes.box_scene(64, seed=seed, law="ISwBSDBF", dt=2e-5, steps=400, omega_max=5.0, spacing=2.02,
                         jitter=0.05, vmax=0.1, offset=3.0, kn=2e3, e=0.6, mu=0.4)
    r = [m for m in w.models if m.name == "Sphere"][0].attributes[0]
    plate = [m for m in w.models if m.name == "Plate"][0]
    thick = plate.attributes[1]
    p = np.array([st.position for st in w.statuses if st.name == "Sphere"])
    w.models.append(DOModel("Lift", plate.group, "orbital", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                            plate.attributes))
    floors = [k for k, st in enumerate(w.statuses) if st.name == "Plate" and tuple(st.orientation_z) == (0, 0, 1)]
    k = min(floors, key=lambda k: w.statuses[k].position[2])
    old = w.statuses[k]
    w.statuses[k] = DOStatus("Lift", (old.position[0], old.position[1], p[:, 2].min() - r - 0.5 * thick - 0.3 * r),
                             (0.0, 0.0, 0.3), old.orientation_x, old.orientation_z, (0.4, -0.3, 5.0))
    sides = [k for k, st in enumerate(w.statuses) if st.name == "Plate" and tuple(st.orientation_z) == (1, 0, 0)]
    k = min(sides, key=lambda k: w.statuses[k].position[0])
    old = w.statuses[k]
    w.statuses[k] = DOStatus("Lift", (p[:, 0].min() - r - 0.5 * thick - 0.2 * r, old.position[1], old.position[2]),
                             (0.2, 0.0, 0.0), old.orientation_x, old.orientation_z, (0.0, 0.0, 0.0))
    return w


def zoi_world(w, vz=-0.4, below=0.35):
    """No floor plate; every sphere moves down at vz; the zone of interest
    ends `below` radii under the lowest sphere centres."""
    from vedo_amd.world import DOStatus
    r = [m for m in w.models if m.name == "Sphere"][0].attributes[0]
    floor = [k for k, st in enumerate(w.statuses) if st.name == "Plate" and tuple(st.orientation_z) == (0, 0, 1)]
    k = min(floor, key=lambda k: w.statuses[k].position[2])
    del w.statuses[k]
    zmin = min(st.position[2] for st in w.statuses if st.name == "Sphere")
    w.statuses = [DOStatus(st.name, st.position, (st.velocity[0], st.velocity[1], st.velocity[2] + vz),
                           st.orientation_x, st.orientation_z, st.angular_velocity)
                  if st.name == "Sphere" else st for st in w.statuses]
    sp = w.system_parameter
    sp.zoi_lower = (0.0, 0.0, zmin - below * r)
    sp.zoi_upper = (0.0, 0.0, 1.0)
    return w


def piston_world(w, density=100.0, force=-0.02, overlap=0.02):
    """The box's top plate as a constrained piston touching the top spheres."""
    from vedo_amd.world import DOModel, DOStatus
    sph = [st for st in w.statuses if st.name == "Sphere"]
    r = [m for m in w.models if m.name == "Sphere"][0].attributes[0]
    plate = [m for m in w.models if m.name == "Plate"][0]
    thick = plate.attributes[1]
    ztop = max(st.position[2] for st in sph)
    piston = DOModel("Piston", plate.group, "constrained", "global", density, 1.0, (0.0, 0.0, force),
                     "QuasiPlate", plate.attributes)
    tops = [k for k, st in enumerate(w.statuses) if st.name == "Plate" and tuple(st.orientation_z) == (0, 0, 1)]
    k = max(tops, key=lambda k: w.statuses[k].position[2])
    old = w.statuses[k]
    w.statuses[k] = DOStatus("Piston", (old.position[0], old.position[1], ztop + r + 0.5 * thick - overlap * r),
                             (0.0, 0.0, 0.0), old.orientation_x, old.orientation_z, (0.0, 0.0, 0.0))
    w.models.append(piston)
    return w


def corner_world(w, density=100.0):
    """The x-min and y-min side plates as constrained walls (no external
    force), each moved to touch the spheres of its side."""
    from vedo_amd.world import DOModel, DOStatus
    sph = [st for st in w.statuses if st.name == "Sphere"]
    r = [m for m in w.models if m.name == "Sphere"][0].attributes[0]
    plate = [m for m in w.models if m.name == "Plate"][0]
    thick = plate.attributes[1]
    for axis, name in ((0, "WallX"), (1, "WallY")):
        sides = [k for k, st in enumerate(w.statuses)
                 if st.name == "Plate" and abs(st.orientation_z[axis]) == 1.0]
        k = min(sides, key=lambda k: w.statuses[k].position[axis])
        old = w.statuses[k]
        lo = min(st.position[axis] for st in sph)
        p = list(old.position)
        p[axis] = lo - r - 0.5 * thick + 0.02 * r
        w.statuses[k] = DOStatus(name, tuple(p), (0.0, 0.0, 0.0), old.orientation_x, old.orientation_z,
                                 (0.0, 0.0, 0.0))
        w.models.append(DOModel(name, plate.group, "constrained", "global", density, 1.0, (0.0, 0.0, 0.0),
                                "QuasiPlate", plate.attributes))
    return w


# knight analyses on the final state of these cases (CalcCoordinateNum,
# CalcDensity): regions from the spheres' final bounding box
METRIC_CASES = ("box_bsdbf", "box_h", "drum_kin", "box_piston")


def metric_regions(state, sphere):
    """(kind, box) requests: the spheres' bounding box, its lower half and its
    middle third, as xmin xmax ymin ymax zmin zmax."""
    p = state[sphere, 0]
    lo, hi = p.min(0), p.max(0)
    whole = np.array([lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]])
    lower = whole.copy()
    lower[5] = 0.5 * (lo[2] + hi[2])
    mid = np.array([lo[0] + (hi[0] - lo[0]) / 3, hi[0] - (hi[0] - lo[0]) / 3, lo[1] + (hi[1] - lo[1]) / 3,
                    hi[1] - (hi[1] - lo[1]) / 3, lo[2] + (hi[2] - lo[2]) / 3, hi[2] - (hi[2] - lo[2]) / 3])
    return [("coord", whole), ("coord", lower), ("coord", mid), ("density", whole), ("density", lower)]


def hole_world(w, hole=0.8, off=(0.1, -0.05), vz=-0.3):
    """The box's floor as a fixed plate with a circular hole
    (DOFixedQuasiPlateWithCircularHole), raised to just under the lowest
    spheres and turned a quarter about z (Ox = +y), so that the hole centre's
    global-axes offset (CDSphere2QuasiPlateWithCircularHole.cpp:46) and the
    rim point's plate-axes offset differ; the spheres move down onto the
    plate, its rim and through the hole."""
    from vedo_amd.world import DOModel, DOStatus
    sph = [k for k, st in enumerate(w.statuses) if st.name == "Sphere"]
    r = [m for m in w.models if m.name == "Sphere"][0].attributes[0]
    plate = [m for m in w.models if m.name == "Plate"][0]
    wdt, thick, lng = plate.attributes
    zlow = min(w.statuses[k].position[2] for k in sph)
    floors = [k for k, st in enumerate(w.statuses) if st.name == "Plate" and tuple(st.orientation_z) == (0, 0, 1)]
    k = min(floors, key=lambda k: w.statuses[k].position[2])
    old = w.statuses[k]
    # a hole smaller than a sphere, a little off the plate's centre (`off`
    # radii), the plate moved so that the hole lies under a sphere of the
    # lowest layer: that sphere rests on the rim.  The rim point is placed
    # along the turned plate's axes, the hole's centre along global ones: the
    # small offset makes them differ
    low = [q for q in sph if w.statuses[q].position[2] < zlow + 0.5 * r]
    q0 = low[len(low) // 2]
    hx, hy = off[0] * r, off[1] * r
    px = w.statuses[q0].position[0] - hx
    py = w.statuses[q0].position[1] - hy
    w.models.append(DOModel("HolePlate", plate.group, "fixed", "global", plate.density, 1.0, (0.0, 0.0, 0.0),
                            "QuasiPlateWithCircularHole", (lng, thick, wdt, hole * r, hx, hy)))
    w.statuses[k] = DOStatus("HolePlate", (px, py, zlow - r - 0.5 * thick + 0.02 * r),
                             (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.0, 0.0, 0.0))
    for q in sph:
        st = w.statuses[q]
        w.statuses[q] = DOStatus(st.name, st.position, (st.velocity[0], st.velocity[1], st.velocity[2] + vz),
                                 st.orientation_x, st.orientation_z, st.angular_velocity)
    return w


def run_case(name, wa, consultant, nsteps, dumps, assigned, outdir=HERE, metrics=None):
    with tempfile.TemporaryDirectory() as td:
        scene = os.path.join(td, "scene.txt")
        write_vscene(wa, scene, consultant)
        args = [HARNESS, scene, os.path.join(td, "out.bin"), str(nsteps), ",".join(map(str, dumps))]
        if metrics is not None:
            with open(os.path.join(td, "metrics.txt"), "w") as f:
                for kind, box in metrics:
                    f.write(kind + " " + " ".join(repr(float(x)) for x in box) + "\n")
            args += ["--metrics", os.path.join(td, "metrics.txt"), os.path.join(td, "metrics_out.txt")]
        if assigned is not None:
            ids, st = assigned
            with open(os.path.join(td, "assigned.bin"), "wb") as f:
                f.write(struct.pack("<i", len(ids)))
                f.write(np.asarray(ids, "<u4").tobytes())
                f.write(np.ascontiguousarray(st, "<f8").tobytes())
            args += ["--assigned", os.path.join(td, "assigned.bin")]
        subprocess.run(args, cwd=td, check=True, stdout=subprocess.DEVNULL)
        d = parse_dumps(os.path.join(td, "out.bin"))
        values = None
        if metrics is not None:
            with open(os.path.join(td, "metrics_out.txt")) as f:
                values = np.array([float(x) for x in f.read().split()])
            assert len(values) == len(metrics)
    arrays = dict(
        consultant=np.int32(consultant), nsteps=np.int32(nsteps), dumps=np.array(dumps, np.int32),
        model_shape=wa.model_shape, model_behavior=wa.model_behavior, model_scope=wa.model_scope,
        model_group=wa.model_group, model_density=wa.model_density,
        model_density_factor=wa.model_density_factor, model_fext=wa.model_fext, model_attr=wa.model_attr,
        iact_law=wa.iact_law, iact_param=wa.iact_param, model_of=wa.model_of, pos=wa.pos, vel=wa.vel,
        ox=wa.ox, oz=wa.oz, av=wa.av,
        scalars=np.array([wa.dt, wa.t_start, wa.t_stop, wa.t_current, wa.safety_factor]),
        g=wa.g, cadence=np.array([wa.up_iact, wa.record], np.int64),
        zoi=np.concatenate([np.asarray(wa.zoi_lo, np.float64), np.asarray(wa.zoi_hi, np.float64)]),
        pbc=np.concatenate([np.asarray(wa.pbc_lo, np.float64), np.asarray(wa.pbc_hi, np.float64)]),
        model_hole=np.asarray(wa.model_hole, np.float64))
    if assigned is not None:
        arrays["assigned_ids"] = np.asarray(assigned[0], np.uint32)
        arrays["assigned"] = assigned[1]
    for k, rec in enumerate(d):
        arrays[f"d{k}_step"] = np.int32(rec["step"])
        arrays[f"d{k}_t"] = np.float64(rec["t"])
        arrays[f"d{k}_state"] = rec["state"]
        arrays[f"d{k}_iact"] = rec["iact"]
        arrays[f"d{k}_status"] = rec["status"]
    if metrics is not None:
        # knight analyses of the final state (after nsteps)
        arrays["metric_kind"] = np.array([0 if k == "coord" else 1 for k, _ in metrics], np.int32)
        arrays["metric_box"] = np.array([b for _, b in metrics], np.float64)
        arrays["metric_value"] = values
    np.savez_compressed(os.path.join(outdir, f"{name}.npz"), **arrays)
    return d


IDO_DUMPS = [1, 2, 6, 7, 8, 30, 60]


def make_ido_restart(outdir=HERE, name="ido_restart", w=None):
    """A restart from the reference's own checkpoint: a lattice pressed a little
    together runs `record` steps and RecordIDO writes ref_run_000001.ido
    (Consultant.cpp:213-229); a second harness process reads that file with
    DOWorld::ReadIDO and runs 60 more steps, writing its own first .ido after
    `record` of them.  Saved: both files' bytes, the cadence and the
    restart's dumps.  `w` replaces the lattice (ido_restart_pbc: a periodic
    box, whose files carry the periodic boundary)."""
    import shutil
    if w is None:
        w = scenes.box_scene(64, seed=18, law="ISwBSDB2F", dt=2e-5, steps=400, omega_max=5.0, spacing=1.99,
                             jitter=0.01, vmax=0.02, offset=0.0, kn=2e3, e=0.6, mu=0.4)
    up, rec, nsteps = 7, 10, 60
    wa = w.to_arrays(up_iact=up, record=rec)
    with tempfile.TemporaryDirectory() as td:
        scene = os.path.join(td, "scene.txt")
        write_vscene(wa, scene, 0)
        subprocess.run([HARNESS, scene, os.path.join(td, "a.bin"), str(rec), str(rec)], cwd=td, check=True,
                       stdout=subprocess.DEVNULL)
        rd = os.path.join(td, "restart")
        os.mkdir(rd)
        shutil.copy(os.path.join(td, "ref_run_000001.ido"), os.path.join(rd, "start.ido"))
        subprocess.run([HARNESS, "start.ido", "out.bin", str(nsteps), ",".join(map(str, IDO_DUMPS)), "--cadence",
                        str(up), str(rec), "0"], cwd=rd, check=True, stdout=subprocess.DEVNULL)
        # DOWorld::WriteXML after `record` restart steps (state and records)
        xd = os.path.join(td, "xml")
        os.mkdir(xd)
        shutil.copy(os.path.join(td, "ref_run_000001.ido"), os.path.join(xd, "start.ido"))
        subprocess.run([HARNESS, "start.ido", "x.bin", str(rec), str(rec), "--cadence", str(up), str(rec), "0",
                        "--xml", "world.xml"], cwd=xd, check=True, stdout=subprocess.DEVNULL)
        with open(os.path.join(xd, "world.xml"), "rb") as f:
            xml = f.read()
        with open(os.path.join(rd, "start.ido"), "rb") as f:
            ido0 = f.read()
        with open(os.path.join(rd, "ref_run_000001.ido"), "rb") as f:
            ido1 = f.read()
        d = parse_dumps(os.path.join(rd, "out.bin"))
    arrays = dict(ido0=np.frombuffer(ido0, np.uint8), ido1=np.frombuffer(ido1, np.uint8),
                  xml=np.frombuffer(xml, np.uint8),
                  cadence=np.array([up, rec], np.int64), nsteps=np.int32(nsteps), dumps=np.array(IDO_DUMPS, np.int32))
    for k, r in enumerate(d):
        arrays[f"d{k}_step"] = np.int32(r["step"])
        arrays[f"d{k}_t"] = np.float64(r["t"])
        arrays[f"d{k}_state"] = r["state"]
        arrays[f"d{k}_iact"] = r["iact"]
        arrays[f"d{k}_status"] = r["status"]
    np.savez_compressed(os.path.join(outdir, f"{name}.npz"), **arrays)
    print(f"{name}: {len(ido0)} + {len(ido1)} bytes of .ido, active {[int(r['iact']['active'].sum()) for r in d]}")


def main():
    """Writes every case, or only those named on the command line."""
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    only = set(sys.argv[1:])
    if not only or "ido_restart" in only:
        make_ido_restart()
    if not only or "ido_restart_pbc" in only:
        make_ido_restart(name="ido_restart_pbc",
                         w=scenes.periodic_box((6, 6, 4), "xy", seed=34, law="ISwBSDBF", vmax=0.6))
    for name, wa, cons, nsteps, dumps, assigned in make_cases():
        if only and name not in only:
            continue
        d = run_case(name, wa, cons, nsteps, dumps, assigned)
        if name in METRIC_CASES:
            assert dumps[-1] == nsteps
            sphere = wa.model_shape[wa.model_of] == 1
            req = metric_regions(d[-1]["state"], sphere)
            d = run_case(name, wa, cons, nsteps, dumps, assigned, metrics=req)
        nact = [int(r["iact"]["active"].sum()) for r in d]
        print(f"{name}: {wa.n_do} DOs, {len(d)} dumps, interactions {[len(r['iact']) for r in d]}, active {nact}")


if __name__ == "__main__":
    main()
