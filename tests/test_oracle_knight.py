"""Knight analyses in the oracle (src/vedo/knight): coordination number and
packing density against the compiled reference's CalcCoordinateNum /
CalcDensity on the final states of four fixtures (bit for bit, the density
through njr::RandomGenerator's sequence); granular temperature
(CalculateGranularTemperature, Knight.cpp:1162-1225, whose source file needs
the Boost-based XML reader and so is not compiled here) against closed forms
and a literal restatement of its loops."""
import dataclasses
import math

import numpy as np
import pytest

from conftest import load_golden, step_golden

METRIC_CASES = ["box_bsdbf", "box_h", "drum_kin", "box_piston"]


@pytest.mark.parametrize("name", METRIC_CASES)
def test_knight_metrics_match_reference(name):
    from oracle.oracle import OracleSim
    z, wa = load_golden(name)
    o = OracleSim(wa)
    for k in range(int(z["nsteps"])):
        step_golden(o, z, k)
    for kind, box, ref in zip(z["metric_kind"], z["metric_box"], z["metric_value"]):
        v = o.coordination_number(box) if kind == 0 else o.packing_density(box)
        assert v == ref or (math.isnan(v) and math.isnan(ref)), (kind, box, v, ref)


def _lattice(n_side, r, spacing, vel=None, av=None):
    """Spheres on a cubic lattice of the given spacing from a dyadic origin (with
    dyadic r and spacing every coordinate and distance is exact)."""
    from vedo_amd import scenes
    n = n_side ** 3
    wa = scenes.box_arrays(n, seed=3, radius=r, law="ISwBSDBF", steps=10)
    sph = np.nonzero(wa.model_shape[wa.model_of] == 1)[0]
    g = np.stack(np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = wa.pos.copy()
    pos[sph] = 0.0625 + spacing * g[:len(sph)]
    out = dataclasses.replace(wa, pos=pos)
    if vel is not None:
        v = wa.vel.copy(); v[sph] = vel; out = dataclasses.replace(out, vel=v)
    if av is not None:
        w = wa.av.copy(); w[sph] = av; out = dataclasses.replace(out, av=w)
    return out, sph


def test_coordination_number_of_a_touching_lattice():
    from oracle.oracle import OracleSim
    r = 2.0 ** -10
    wa, sph = _lattice(5, r, 2 * r)
    o = OracleSim(wa)
    p = wa.pos[sph]
    lo, hi = p.min(0), p.max(0)
    # the interior 3x3x3 spheres only (the box widened by r must not reach the outer layer)
    inner = np.array([lo[0] + 1.6 * r, hi[0] - 1.6 * r, lo[1] + 1.6 * r, hi[1] - 1.6 * r, lo[2] + 1.6 * r,
                      hi[2] - 1.6 * r])
    assert o.coordination_number(inner) == 6.0
    # all 125: 8 corners (3), 36 edges (4), 54 faces (5), 27 interior (6)
    whole = np.array([lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]])
    assert o.coordination_number(whole) == (8 * 3 + 36 * 4 + 54 * 5 + 27 * 6) / 125
    # apart by more than 2r: none
    wa2, sph2 = _lattice(3, r, 2 * r + 2.0 ** -20)
    assert OracleSim(wa2).coordination_number(whole * 0 + [0, 1, 0, 1, 0, 1]) == 0.0


def test_packing_density_limits():
    from oracle.oracle import OracleSim
    r = 2.0 ** -10
    wa, sph = _lattice(4, r, 2 * r)
    o = OracleSim(wa)
    p = wa.pos[sph]
    # a box inside one sphere: every point hits
    c = p[21]
    assert o.packing_density([c[0] - 0.3 * r, c[0] + 0.3 * r, c[1] - 0.3 * r, c[1] + 0.3 * r, c[2] - 0.3 * r,
                              c[2] + 0.3 * r], 2000) == 100.0
    # a box far from every sphere: none
    assert o.packing_density([1.0, 1.1, 1.0, 1.1, 1.0, 1.1], 2000) == 0.0
    # a simple cubic packing: pi/6 of the volume, within Monte Carlo error
    lo, hi = p.min(0), p.max(0)
    d = o.packing_density([lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]], 20000)
    assert abs(d - 100 * math.pi / 6) < 2.0


def _gt_literal(wa, model, c, ox, oz, size):
    """CalculateGranularTemperature's loops, statement by statement."""
    def dot(a, b):
        return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]
    oy = (oz[1] * ox[2] - oz[2] * ox[1], oz[2] * ox[0] - oz[0] * ox[2], oz[0] * ox[1] - oz[1] * ox[0])
    av, aw, n, sel = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 0, []
    for i in range(wa.n_do):
        if wa.model_of[i] != model:
            continue
        lp = [float(wa.pos[i, k]) - c[k] for k in range(3)]
        if abs(dot(lp, ox)) <= 0.5 * size[0] and abs(dot(lp, oy)) <= 0.5 * size[1] and abs(dot(lp, oz)) <= 0.5 * size[2]:
            av = [av[k] + float(wa.vel[i, k]) for k in range(3)]
            aw = [aw[k] + float(wa.av[i, k]) for k in range(3)]
            sel.append(i)
            n += 1
    if n:
        av = [1.0 / n * x for x in av]
        aw = [1.0 / n * x for x in aw]
    tv = tw = 0.0
    for i in sel:
        dv = [float(wa.vel[i, k]) - av[k] for k in range(3)]
        dw = [float(wa.av[i, k]) - aw[k] for k in range(3)]
        tv += dot(dv, dv)
        tw += dot(dw, dw)
    if n:
        tv = tv / n / 3.0
        tw = tw / n / 3.0
    return np.array([n, math.sqrt(dot(av, av)), tv, math.sqrt(dot(aw, aw)), tw])


def test_granular_temperature_closed_forms():
    from oracle.oracle import OracleSim
    r = 1e-3
    rng = np.random.default_rng(7)
    # equal (dyadic) velocities: zero temperature, mean = that velocity
    wa, sph = _lattice(4, r, 2.5 * r, vel=np.array([0.25, -0.5, 0.0]), av=np.array([0.0, 0.0, 2.0]))
    p = wa.pos[sph]
    c = 0.5 * (p.min(0) + p.max(0))
    size = p.max(0) - p.min(0) + 2 * r
    m = int(wa.model_of[sph[0]])
    out = OracleSim(wa).granular_temperature(m, c, size)
    assert out[0] == len(sph) and out[1] == math.sqrt(0.3125) and out[2] == 0.0 and out[3] == 2.0 and out[4] == 0.0
    # +-u on alternate spheres: mean 0, T = |u|^2 / 3
    u = np.array([0.2, 0.1, -0.3])
    vel = np.where((np.arange(len(sph)) % 2 == 0)[:, None], u, -u)
    wa, sph = _lattice(4, r, 2.5 * r, vel=vel)
    out = OracleSim(wa).granular_temperature(m, c, size)
    assert out[0] == len(sph) and out[1] == 0.0
    assert abs(out[2] - u @ u / 3) <= 1e-15
    # random velocities, a rotated rectangle holding part of the lattice: the loops literally
    vel = rng.normal(size=(len(sph), 3))
    av = rng.normal(size=(len(sph), 3))
    wa, sph = _lattice(4, r, 2.5 * r, vel=vel, av=av)
    a = 0.4
    ox = (math.cos(a), math.sin(a), 0.0)
    oz = (0.0, 0.0, 1.0)
    sz = (0.6 * size[0], 0.5 * size[1], 0.7 * size[2])
    out = OracleSim(wa).granular_temperature(m, c, sz, ox=ox, oz=oz)
    ref = _gt_literal(wa, m, c, ox, oz, sz)
    assert 0 < out[0] < len(sph)
    assert np.array_equal(out, ref), (out, ref)
    # another model selects nothing
    out = OracleSim(wa).granular_temperature(m + 1, c, size)
    assert out[0] == 0 and out[2] == 0.0
