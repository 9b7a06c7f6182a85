"""Seeded synthetic scenes for the DEM step (box settling, rotating drum).

Every random number comes from the project's own splitmix64 stream with
53-bit uniforms and Box-Muller normals, so any host regenerates a scene
bit-for-bit.  The scenes are described with the DOWorld mirror in world.py,
using the reference's model/mechanism names.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

from .world import DOModel, DOStatus, DOWorld, IactModel, SystemParameter

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with vectorised draws."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & MASK64)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            inc = np.uint64(0x9E3779B97F4A7C15)
            z = self.state + inc * np.arange(1, n + 1, dtype=np.uint64)
            self.state = self.state + inc * np.uint64(n)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2)])
        return z[:n]


# ---------------------------------------------------------------------------
# Contact-law parameter maps
# ---------------------------------------------------------------------------

def damping_ratio(e: float) -> float:
    """beta = -ln e / sqrt(pi^2 + ln^2 e), the e->beta map of
    ISwModels::WetDampingRatioLegendre2006 (ISwModels.cpp:282-290)."""
    le = math.log(e)
    return -le / math.sqrt(math.pi * math.pi + le * le)


def hertz_mindlin(E: float, nu: float, mu: float, e: float, r_a: float, r_b: Optional[float],
                  m_a: float, m_b: Optional[float]) -> dict:
    """ISwH mechanisms for two spheres (r_b, m_b given) or a sphere against a
    rigid wall (r_b, m_b None): k_H = 4/3 E* sqrt(R*), EquivalentMass = M*,
    ShearStiffnessCoefficient1 = 8 G*, ShearStiffnessCoefficient2 = 3 R*/(4 E*)
    so that ks = 8 G* sqrt(R* delta) under the Hertz force-overlap relation."""
    G = E / (2.0 * (1.0 + nu))
    if r_b is None:
        e_star = E / (1.0 - nu * nu)
        g_star = G / (2.0 - nu)
        r_star = r_a
        m_star = m_a
    else:
        e_star = E / (2.0 * (1.0 - nu * nu))
        g_star = G / (2.0 * (2.0 - nu))
        r_star = r_a * r_b / (r_a + r_b)
        m_star = m_a * m_b / (m_a + m_b)
    return {
        "HertzNormalStiffness": 4.0 / 3.0 * e_star * math.sqrt(r_star),
        "EquivalentMass": m_star,
        "ShearStiffnessCoefficient1": 8.0 * g_star,
        "ShearStiffnessCoefficient2": 3.0 * r_star / (4.0 * e_star),
        "DampingRatio": damping_ratio(e),
        "KineticFrictionCoefficient": mu,
    }


def spring_dashpot(kn: float, kt_ratio: float, e: float, mu: float, equal_spheres: bool = True) -> dict:
    """ISwBSDBF mechanisms.  ISwModels::CriticalDamping uses the arithmetic
    mean of the two masses (ISwModels.cpp:88-108), so for two equal spheres the
    damping ratio that yields restitution e is zeta/sqrt(2)."""
    beta = damping_ratio(e) / (math.sqrt(2.0) if equal_spheres else 1.0)
    return {
        "NormalStiffness": kn, "ShearStiffness": kt_ratio * kn,
        "NormalDampingRatio": beta, "ShearDampingRatio": beta,
        "NormalBondStrength": 0.0, "ShearBondStrength": 0.0, "FrictionCoefficient": mu,
    }


def _plate(name: str, center, ox, oz) -> DOStatus:
    return DOStatus(name, tuple(center), (0, 0, 0), tuple(ox), tuple(oz), (0, 0, 0))


def box_walls(side: float, thick: float, group: str = "Wall") -> Tuple[List[DOModel], List[DOStatus]]:
    """Six fixed quasi-plates whose mid-planes are the faces of [0,side]^3."""
    w = side + 2.0 * thick
    m = DOModel("Plate", group, "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate", (w, thick, w))
    c = 0.5 * side
    st = [
        _plate("Plate", (c, c, 0.0), (1, 0, 0), (0, 0, 1)),
        _plate("Plate", (c, c, side), (1, 0, 0), (0, 0, 1)),
        _plate("Plate", (0.0, c, c), (0, 1, 0), (1, 0, 0)),
        _plate("Plate", (side, c, c), (0, 1, 0), (1, 0, 0)),
        _plate("Plate", (c, 0.0, c), (0, 0, 1), (0, 1, 0)),
        _plate("Plate", (c, side, c), (0, 0, 1), (0, 1, 0)),
    ]
    return [m], st


def box_scene(n: int, seed: int = 1, radius: float = 1e-3, density: float = 2500.0,
              side: float = 0.1, law: str = "ISwH", E: float = 1e7, nu: float = 0.3, mu: float = 0.5,
              e: float = 0.5, kn: float = 1e4, dt: float = 1e-6, steps: int = 2000,
              spacing: float = 2.5, jitter: float = 0.1, vmax: float = 0.1, thick: float = 2e-3,
              walls: bool = True, offset: float = 5.0, gravity: float = -9.81,
              omega_max: float = 0.0) -> DOWorld:
    """Monodisperse spheres on a seeded jittered cubic lattice in a closed box.

    Lattice spacing ``spacing*r``, x fastest, first ``n`` sites of the
    smallest near-cubic block, offset ``offset*r`` from the box corner;
    jitter U[-jitter*r, jitter*r] per axis; velocity U[-vmax, vmax] per axis."""
    rng = SplitMix64(seed)
    s = spacing * radius
    nx = max(1, int(math.ceil(n ** (1.0 / 3.0))))
    ny = nx
    nz = max(1, int(math.ceil(n / (nx * ny))))
    idx = np.arange(n)
    ix, iy, iz = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    base = offset * radius + np.stack([ix, iy, iz], 1) * s
    pos = base + rng.uniform(3 * n, -jitter * radius, jitter * radius).reshape(n, 3)
    vel = rng.uniform(3 * n, -vmax, vmax).reshape(n, 3)
    av = rng.uniform(3 * n, -omega_max, omega_max).reshape(n, 3) if omega_max else np.zeros((n, 3))
    side = max(side, float(pos.max()) + offset * radius)
    sphere = DOModel("Sphere", "Granule", "mobile", "local", density, 1.0, (0, 0, 0), "Sphere", (radius,))
    mass = sphere.analysis()[1]
    models = [sphere]
    statuses = [DOStatus("Sphere", tuple(pos[i]), tuple(vel[i]), (1, 0, 0), (0, 0, 1), tuple(av[i]))
                for i in range(n)]
    iact = []
    if law == "ISwH":
        iact.append(IactModel("Granule", "Granule", "ISwH", hertz_mindlin(E, nu, mu, e, radius, radius, mass, mass)))
        wall_mech = hertz_mindlin(E, nu, mu, e, radius, None, mass, None)
    elif law in ("ISwBSDBF", "ISwBSDB2F", "ISwBtSDBF"):
        iact.append(IactModel("Granule", "Granule", law, spring_dashpot(kn, 0.8, e, mu, True)))
        wall_mech = spring_dashpot(kn, 0.8, e, mu, False)
    elif law == "ISwHertz":
        k_h = hertz_mindlin(E, nu, mu, e, radius, radius, mass, mass)["HertzNormalStiffness"]
        iact.append(IactModel("Granule", "Granule", "ISwHertz",
                              {"NormalStiffness": k_h, "NormalDampingCoefficient": 1e-3}))
        wall_mech = {"NormalStiffness": 2.0 * k_h, "NormalDampingCoefficient": 1e-3}
    else:
        raise ValueError(law)
    if walls:
        wm, ws = box_walls(side, thick)
        models += wm
        statuses += ws
        iact.append(IactModel("Wall", "Granule", law, wall_mech))
    sp = SystemParameter(0.0, steps * dt, dt, 0.0, (0.0, 0.0, gravity))
    return DOWorld(sp, models, iact, statuses)


def lognormal_classes(k: int, median: float, sigma: float, rmin: float, rmax: float) -> np.ndarray:
    """K log-spaced radius classes spanning the clipped range."""
    return np.exp(np.linspace(math.log(rmin), math.log(rmax), k))


def drum_scene(n: int, seed: int = 2, median: float = 1e-3, sigma: float = 0.2, rmin: float = 0.5e-3,
               rmax: float = 2e-3, classes: int = 64, density: float = 2500.0, R: float = 0.15,
               L: Optional[float] = None, omega: float = 2.0, kn: float = 1e4, kt_ratio: float = 0.8,
               e: float = 0.6, mu: float = 0.4, dt: float = 5e-6, steps: int = 10000,
               sides: int = 64, thick: float = 2e-3, gap: float = 1e-6, arrays: bool = False,
               up_iact: int = 20, record: int = 10 ** 9):
    """Polydisperse spheres in a rotating drum of quasi-plates.

    Radii are log-normal (median, sigma) clipped to [rmin, rmax] and quantised
    to ``classes`` log-spaced classes, one DOModel each in one group.  The drum
    axis is y; its mantle is a ``sides``-gon of quasi-plates plus two end
    plates, all prescribed to rotate rigidly at ``omega`` (kinematic walls).
    Spheres start as a settled bed in the lower half of the drum: per radius
    class a slab of triangular lattices in the cross-section, pocket contacts
    pre-compressed to the equilibrium of their load, in-row neighbours
    2*(1+gap)*r apart.  When ``L`` is None the drum is made just long enough
    for that bed.

    Returns (world, wall_ids, axis_point, axis_dir); with ``arrays=True`` the
    first item is a WorldArrays built without per-sphere Python objects (for
    million-sphere scenes)."""
    rng = SplitMix64(seed)
    cls = lognormal_classes(classes, median, sigma, rmin, rmax)
    r_raw = np.clip(median * np.exp(sigma * rng.normal(n)), rmin, rmax)
    edges = np.sqrt(cls[1:] * cls[:-1])
    ci = np.searchsorted(edges, r_raw)
    radii = cls[ci]
    # Sort by class so that each lattice layer holds similar radii.
    order = np.argsort(ci, kind="stable")
    ci_sorted = ci[order]
    # Fill of the lower half of the drum, one slab per radius class along the
    # axis (y).  Each y-layer of a slab is a triangular lattice of equal
    # spheres in the cross-section: rows at spacing a = 2r(1+gap), alternate
    # rows offset by a/2, so that every sphere sits in the pocket of two
    # spheres of the row below.  The pocket contacts are pre-compressed to
    # the linear-spring equilibrium of the weight above them (W = rows above
    # * m g shared by two contacts at 30 degrees from vertical), so the bed
    # starts at rest and stays so; in-row and y neighbours start a hair
    # apart and carry no load.
    margin = thick * 0.5 + 1e-9
    rin = R - margin
    g = 9.81
    cos30 = math.sqrt(3.0) / 2.0

    def rows_z(r, a, M):
        lim = rin - r
        mg = density * 4.0 / 3.0 * math.pi * r ** 3 * g
        z = [-lim]
        for j in range(M - 1):
            d = 2.0 * r - (M - 1 - j) * mg / (2.0 * cos30 * kn)
            z.append(z[-1] + math.sqrt(d * d - 0.25 * a * a))
        return lim, z

    def section(r, a, M):
        lim, zr = rows_z(r, a, M)
        xs, zs = [], []
        for m, z in enumerate(zr):
            if z > 0.0:
                break
            half = math.sqrt(max(0.0, lim * lim - z * z))
            k = np.arange(-int(half / a) - 1, int(half / a) + 2)
            x = (k + 0.5 * (m % 2)) * a
            x = x[np.abs(x) <= half]
            xs.append(x)
            zs.append(np.full(len(x), z))
        X, Z = np.concatenate(xs), np.concatenate(zs)
        o = np.lexsort((X, Z))
        return X[o], Z[o]

    pos = np.zeros((n, 3))
    y = margin
    placed = 0
    for c in np.unique(ci_sorted):
        m = int(np.sum(ci_sorted == c))
        r = float(cls[c])
        a = 2.0 * (1.0 + gap) * r
        m_full = int((rin - r) / math.sqrt(3.0 * r * r)) + 1
        X, Z = section(r, a, m_full)
        cap = len(X)
        left = m
        while left > 0:
            take = min(cap, left)
            if take < cap:
                # last, partial layer: as many rows as it fills, loaded by those
                lo, hi = 1, m_full
                while lo < hi:
                    mid = (lo + hi) // 2
                    if len(section(r, a, mid)[0]) >= take:
                        hi = mid
                    else:
                        lo = mid + 1
                X, Z = section(r, a, lo)
            sl = slice(placed, placed + take)
            pos[sl, 0] = X[:take]
            pos[sl, 1] = y + r
            pos[sl, 2] = Z[:take]
            placed += take
            left -= take
            y += a
        y += 1e-3 * rmin
    Ld = y + margin
    if L is not None:
        if Ld > L:
            raise ValueError("drum too short for the requested sphere count")
        Ld = L
    # undo the class sort: sphere i keeps its own radius
    pos_final = np.empty_like(pos)
    pos_final[order] = pos
    jitter = rng.uniform(3 * n, -1e-8, 1e-8).reshape(n, 3) * radii[:, None]
    pos_final += jitter
    models = []
    for k, r in enumerate(cls):
        models.append(DOModel(f"S{k:02d}", "Granule", "mobile", "local", density, 1.0, (0, 0, 0),
                              "Sphere", (float(r),)))
    statuses = [] if arrays else [DOStatus(f"S{ci[i]:02d}", tuple(pos_final[i])) for i in range(n)]
    # drum walls: mantle plates tangent to the circle of radius R, axis along y
    chord = 2.0 * R * math.sin(math.pi / sides) * 1.02
    models.append(DOModel("Mantle", "Wall", "orbital", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                          (chord, thick, Ld + 2.0 * thick)))
    models.append(DOModel("EndCap", "Wall", "orbital", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                          (2.0 * R + 4.0 * thick, thick, 2.0 * R + 4.0 * thick)))
    wall_ids = []
    for k in range(sides):
        th = 2.0 * math.pi * k / sides           # plate 0 is the bottom of the drum
        c = (R * math.sin(th), 0.5 * Ld, -R * math.cos(th))
        oz = (math.sin(th), 0.0, -math.cos(th))         # outward normal
        ox = (math.cos(th), 0.0, math.sin(th))          # tangential
        v = (omega * c[2], 0.0, -omega * c[0])          # omega about +y: w x r
        wall_ids.append(n + len(wall_ids))
        statuses.append(DOStatus("Mantle", c, v, ox, oz, (0.0, omega, 0.0)))
    for y in (0.0, Ld):
        wall_ids.append(n + len(wall_ids))
        statuses.append(DOStatus("EndCap", (0.0, y, 0.0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0.0, omega, 0.0)))
    iact = [IactModel("Granule", "Granule", "ISwBSDBF", spring_dashpot(kn, kt_ratio, e, mu, True)),
            IactModel("Wall", "Granule", "ISwBSDBF", spring_dashpot(kn, kt_ratio, e, mu, False))]
    sp = SystemParameter(0.0, steps * dt, dt, 0.0, (0.0, 0.0, -9.81))
    world = DOWorld(sp, models, iact, statuses)
    if arrays:
        wa = world.to_arrays(up_iact, record)
        wall_model = wa.model_of
        one = np.ones((n, 1))
        wa.model_of = np.concatenate([ci.astype(np.int32), wall_model])
        wa.pos = np.concatenate([pos_final, wa.pos])
        wa.vel = np.concatenate([np.zeros((n, 3)), wa.vel])
        wa.ox = np.concatenate([one * np.array([1.0, 0.0, 0.0]), wa.ox])
        wa.oz = np.concatenate([one * np.array([0.0, 0.0, 1.0]), wa.oz])
        wa.av = np.concatenate([np.zeros((n, 3)), wa.av])
        return wa, wall_ids, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)
    return world, wall_ids, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)


def rigid_rotation_status(pos0, vel0, ox0, oz0, axis_point, axis_dir, omega: float, t: float):
    """Statuses of walls rotating rigidly about an axis by angle omega*t
    (Rodrigues' rotation), with velocity omega x (x - c) and angular
    velocity omega*axis.  Inputs are [W,3] arrays of the t=0 statuses."""
    a = np.asarray(axis_dir, np.float64)
    a = a / np.linalg.norm(a)
    c = np.asarray(axis_point, np.float64)
    th = omega * t
    ct, s = math.cos(th), math.sin(th)

    def rot(v):
        return v * ct + np.cross(a, v) * s + np.outer(v @ a, a) * (1.0 - ct)

    p = c + rot(np.asarray(pos0) - c)
    w = omega * a
    v = np.cross(np.broadcast_to(w, p.shape), p - c)
    return p, v, rot(np.asarray(ox0)), rot(np.asarray(oz0)), np.broadcast_to(w, p.shape).copy()
