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

from .world import DOModel, DOStatus, DOWorld, IactModel, SystemParameter, WorldArrays

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


def periodic_box(counts=(6, 6, 4), axes: str = "xy", seed: int = 31, radius: float = 1e-3,
                 density: float = 2500.0, law: str = "ISwBSDBF", spacing: float = 2.02, jitter: float = 0.05,
                 vmax: float = 0.3, omega_max: float = 5.0, dt: float = 2e-5, steps: int = 400,
                 thick: float = 2e-3, gravity: Optional[float] = None, drift=(0.0, 0.0, 0.0),
                 **law_args) -> DOWorld:
    """Spheres on a jittered cubic lattice that fills a periodic cell.

    ``counts`` lattice sites per axis at ``spacing*r``; along each axis in
    ``axes`` the cell is periodic (PeriodicBoundaryConditions from 0 to
    counts*spacing*r, so the first and last layers are one spacing apart
    across the boundary); along every other axis two fixed quasi-plates bound
    the lattice one radius beyond its outer layers, their centres inside the
    periodic range and their sides 1.5 times the cell.  Gravity along -z
    (default -9.81 unless z is periodic, then 0); ``drift`` is added to every
    sphere's velocity (the bed streams through the periodic faces)."""
    rng = SplitMix64(seed)
    s = spacing * radius
    cnt = [int(c) for c in counts]
    per = [a in axes for a in "xyz"]
    n = cnt[0] * cnt[1] * cnt[2]
    idx = np.arange(n)
    ijk = np.stack([idx % cnt[0], (idx // cnt[0]) % cnt[1], idx // (cnt[0] * cnt[1])], 1)
    pos = (ijk + 0.5) * s + rng.uniform(3 * n, -jitter * radius, jitter * radius).reshape(n, 3)
    vel = rng.uniform(3 * n, -vmax, vmax).reshape(n, 3) + np.asarray(drift, np.float64)
    av = rng.uniform(3 * n, -omega_max, omega_max).reshape(n, 3) if omega_max else np.zeros((n, 3))
    ext = [c * s for c in cnt]
    sphere = DOModel("Sphere", "Granule", "mobile", "local", density, 1.0, (0, 0, 0), "Sphere", (radius,))
    mass = sphere.analysis()[1]
    models = [sphere]
    statuses = [DOStatus("Sphere", tuple(pos[i]), tuple(vel[i]), (1, 0, 0), (0, 0, 1), tuple(av[i]))
                for i in range(n)]
    if law == "ISwH":
        mech = hertz_mindlin(law_args.get("E", 1e6), law_args.get("nu", 0.3), law_args.get("mu", 0.5),
                             law_args.get("e", 0.5), radius, radius, mass, mass)
        wall_mech = hertz_mindlin(law_args.get("E", 1e6), law_args.get("nu", 0.3), law_args.get("mu", 0.5),
                                  law_args.get("e", 0.5), radius, None, mass, None)
    elif law in ("ISwBSDBF", "ISwBSDB2F", "ISwBtSDBF"):
        mech = spring_dashpot(law_args.get("kn", 2e3), 0.8, law_args.get("e", 0.6), law_args.get("mu", 0.4), True)
        wall_mech = spring_dashpot(law_args.get("kn", 2e3), 0.8, law_args.get("e", 0.6),
                                   law_args.get("mu", 0.4), False)
    elif law == "ISwHertz":
        k_h = hertz_mindlin(law_args.get("E", 1e6), law_args.get("nu", 0.3), 0.5, 0.5, radius, radius, mass,
                            mass)["HertzNormalStiffness"]
        mech = {"NormalStiffness": k_h, "NormalDampingCoefficient": 1e-3}
        wall_mech = {"NormalStiffness": 2.0 * k_h, "NormalDampingCoefficient": 1e-3}
    else:
        raise ValueError(law)
    iact = [IactModel("Granule", "Granule", law, mech)]
    walled = [d for d in range(3) if not per[d]]
    if walled:
        # plate d: normal along axis d; Ox along the next axis
        size = [1.5 * e + 2.0 * thick for e in ext]
        for d in walled:
            u, v = (d + 1) % 3, (d + 2) % 3
            m = DOModel(f"Wall{'xyz'[d]}", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                        (size[u], thick, size[v]))
            models.append(m)
            ox = [0.0, 0.0, 0.0]
            ox[u] = 1.0
            for side in (0, 1):
                c = [0.5 * e for e in ext]
                c[d] = (-radius - 0.5 * thick) if side == 0 else ext[d] + radius + 0.5 * thick
                oz = [0.0, 0.0, 0.0]
                oz[d] = 1.0 if side == 0 else -1.0
                statuses.append(_plate(m.name, c, ox, oz))
        iact.append(IactModel("Wall", "Granule", law, wall_mech))
    g = gravity if gravity is not None else (0.0 if per[2] else -9.81)
    sp = SystemParameter(0.0, steps * dt, dt, 0.0, (0.0, 0.0, g))
    sp.pbc_lower = (0.0, 0.0, 0.0)
    sp.pbc_upper = tuple(ext[d] if per[d] else 0.0 for d in range(3))
    return DOWorld(sp, models, iact, statuses)


def lognormal_classes(k: int, median: float, sigma: float, rmin: float, rmax: float) -> np.ndarray:
    """K log-spaced radius classes spanning the clipped range."""
    return np.exp(np.linspace(math.log(rmin), math.log(rmax), k))


def drum_scene(n: int, seed: int = 2, median: float = 1e-3, sigma: float = 0.2, rmin: float = 0.5e-3,
               rmax: float = 2e-3, classes: int = 64, density: float = 2500.0, R: float = 0.15,
               L: Optional[float] = None, omega: float = 2.0, kn: float = 1e4, kt_ratio: float = 0.8,
               e: float = 0.6, mu: float = 0.4, dt: float = 5e-6, steps: int = 10000,
               sides: int = 64, thick: float = 2e-3, gap: float = 1e-6, arrays: bool = False,
               up_iact: int = 20, record: int = 10 ** 9, law: str = "ISwBSDBF", E: float = 5e7,
               nu: float = 0.25):
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

    # ISwH gives a contact that starts at rest its full Hertz stiffness
    # (ISwH.cpp:90-103, kn = k_H when |vn| = 0): its bed starts a hair apart
    # instead, so contacts form with a nonzero approach speed
    compress = law != "ISwH"

    def rows_z(r, a, M):
        lim = rin - r
        mg = density * 4.0 / 3.0 * math.pi * r ** 3 * g
        z = [-lim]
        for j in range(M - 1):
            d = 2.0 * r - (M - 1 - j) * mg / (2.0 * cos30 * kn) if compress else a
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
    if law == "ISwH":
        # one IactModel per group pair: Hertz-Mindlin mechanisms of the median
        # sphere; dt from the lightest pair's first-contact stiffness at 1 m/s
        m_med = density * 4.0 / 3.0 * math.pi * median ** 3
        pair = hertz_mindlin(E, nu, mu, e, median, median, m_med, m_med)
        iact = [IactModel("Granule", "Granule", "ISwH", pair),
                IactModel("Wall", "Granule", "ISwH", hertz_mindlin(E, nu, mu, e, median, None, m_med, None))]
        m_min = density * 4.0 / 3.0 * math.pi * float(cls.min()) ** 3
        k = ish_first_contact_kn(pair["HertzNormalStiffness"], pair["EquivalentMass"], pair["DampingRatio"], 1.0)
        dt = critical_dt(tangential_mass(0.5 * m_min), k)
    else:
        iact = [IactModel("Granule", "Granule", law, spring_dashpot(kn, kt_ratio, e, mu, True)),
                IactModel("Wall", "Granule", law, spring_dashpot(kn, kt_ratio, e, mu, False))]
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


# ---------------------------------------------------------------------------
# Array-built scenes for batched and served runs (no per-sphere Python objects)
# ---------------------------------------------------------------------------

def _spliced(world: DOWorld, sphere_model: int, pos, vel, av, up_iact: int, record: int) -> WorldArrays:
    """WorldArrays of `world` (walls only) with n spheres of model index
    `sphere_model` placed first (DO ids 0..n-1), walls after them."""
    wa = world.to_arrays(up_iact, record)
    n = len(pos)
    one = np.ones((n, 1))
    wa.model_of = np.concatenate([np.full(n, sphere_model, np.int32), wa.model_of])
    wa.pos = np.concatenate([np.asarray(pos, np.float64), wa.pos])
    wa.vel = np.concatenate([np.asarray(vel, np.float64), wa.vel])
    wa.ox = np.concatenate([one * np.array([1.0, 0.0, 0.0]), wa.ox])
    wa.oz = np.concatenate([one * np.array([0.0, 0.0, 1.0]), wa.oz])
    wa.av = np.concatenate([np.asarray(av, np.float64), wa.av])
    return wa


def ish_first_contact_kn(k_h: float, m_eff: float, beta: float, vn: float) -> float:
    """ISwH's normal stiffness at first contact for an impact speed vn
    (ISwH.cpp:90-103, the beta < 0.99 branch): the stiffness the critical
    time step of an ISwH scene is based on."""
    return 0.957832746294131 / (1.0 - beta * beta) * m_eff ** 0.2 * k_h ** 0.8 * vn ** 0.4


def critical_dt(m_eff: float, k: float, fraction: float = 0.1) -> float:
    """fraction * ISwModels::CriticalTimeStep = fraction * 2 pi sqrt(m/k)
    (ISwModels.cpp:8-11)."""
    return fraction * 2.0 * math.pi * math.sqrt(m_eff / k)


def tangential_mass(m_eff: float) -> float:
    """Mass of a contact's tangential-rotational mode, the one that bounds the
    stable time step of a law with a tangential spring: a force at the contact
    point of a solid sphere (I = 0.4 m r^2) moves that point as a mass
    m / (1 + m r^2 / I) = 2m/7 would, so a pair of reduced mass M* has
    2/7 M*.  With the normal mode's M* instead, the compiled reference's
    ISwH runs of soft, weakly damped box scenes gain energy until spheres
    cross more than their radius in a step; with this mass they settle."""
    return 2.0 / 7.0 * m_eff


def box_arrays(n: int, seed: int = 1, radius: float = 1e-3, density: float = 2500.0, law: str = "ISwH",
               E: float = 1e7, nu: float = 0.3, mu: float = 0.5, e: float = 0.5, kn: float = 1e4,
               dt: Optional[float] = None, steps: int = 5000, spacing: float = 2.5, jitter: float = 0.1,
               vmax: float = 0.1, thick: float = 2e-3, offset: float = 5.0, gravity: float = -9.81,
               up_iact: int = 20, record: int = 10 ** 9) -> WorldArrays:
    """The box settling scene of the batched configuration: n spheres on a
    seeded jittered cubic lattice (spacing*r, as box_scene) in a closed box
    sized to the lattice (side = lattice extent + offset*r), one law for
    sphere-sphere and sphere-wall.  dt None: 0.1 of the critical time step of
    the sphere-sphere contact's tangential-rotational mode (tangential_mass),
    for ISwH at the impact speed of a fall through the lattice height."""
    rng = SplitMix64(seed)
    s = spacing * radius
    nx = max(1, int(math.ceil(n ** (1.0 / 3.0))))
    nz = max(1, int(math.ceil(n / (nx * nx))))
    idx = np.arange(n)
    ix, iy, iz = idx % nx, (idx // nx) % nx, idx // (nx * nx)
    base = offset * radius + np.stack([ix, iy, iz], 1) * s
    pos = base + rng.uniform(3 * n, -jitter * radius, jitter * radius).reshape(n, 3)
    vel = rng.uniform(3 * n, -vmax, vmax).reshape(n, 3)
    side = float(pos.max()) + offset * radius
    sphere = DOModel("Sphere", "Granule", "mobile", "local", density, 1.0, (0, 0, 0), "Sphere", (radius,))
    mass = sphere.analysis()[1]
    wm, ws = box_walls(side, thick)
    if law == "ISwH":
        pair = hertz_mindlin(E, nu, mu, e, radius, radius, mass, mass)
        wall = hertz_mindlin(E, nu, mu, e, radius, None, mass, None)
        if dt is None:
            v_ref = math.sqrt(2.0 * abs(gravity) * nz * s) + vmax
            k = ish_first_contact_kn(pair["HertzNormalStiffness"], pair["EquivalentMass"],
                                     pair["DampingRatio"], v_ref)
            dt = critical_dt(tangential_mass(pair["EquivalentMass"]), k)
    else:
        pair = spring_dashpot(kn, 0.8, e, mu, True)
        wall = spring_dashpot(kn, 0.8, e, mu, False)
        if dt is None:
            dt = critical_dt(tangential_mass(0.5 * mass), kn)
    iact = [IactModel("Granule", "Granule", law, pair), IactModel("Wall", "Granule", law, wall)]
    world = DOWorld(SystemParameter(0.0, steps * dt, dt, 0.0, (0.0, 0.0, gravity)), [sphere] + wm, iact, ws)
    return _spliced(world, 0, pos, vel, np.zeros((n, 3)), up_iact, record)


def batched_box(k: int, steps: int, up_iact: int = 20, count: int = 256, seed: int = 4) -> WorldArrays:
    """Scene k of the batched configuration: `count` box-settling scenes
    (ISwH), sphere counts log-uniform in [1e3, 2e5], mu ~ U[0.2, 0.7],
    e ~ U[0.3, 0.9], E ~ logU[1e6, 1e8], r = 1 mm, box sized to its lattice,
    the box_arrays time step."""
    rng = SplitMix64(seed)
    u = rng.uniform(4 * count, 0.0, 1.0).reshape(count, 4)
    n = int(round(math.exp(math.log(1e3) + u[k, 0] * (math.log(2e5) - math.log(1e3)))))
    return box_arrays(n, seed=seed * 1000 + k, law="ISwH", mu=0.2 + 0.5 * u[k, 1], e=0.3 + 0.6 * u[k, 2],
                      E=10.0 ** (6.0 + 2.0 * u[k, 3]), steps=steps, up_iact=up_iact)


def hopper_arrays(n: int, seed: int = 3, radius: float = 1e-3, density: float = 2500.0, law: str = "ISwH",
                  E: float = 5e7, nu: float = 0.25, mu: float = 0.5, e: float = 0.5, kn: float = 1e4,
                  outlet: float = 0.03, dt: Optional[float] = None, steps: int = 5000, thick: float = 2e-3,
                  up_iact: int = 20, record: int = 10 ** 9) -> WorldArrays:
    """Spheres discharging from a square hopper: four quasi-plates inclined at
    45 degrees down to a square outlet (side `outlet`, at most half the
    hopper's), four vertical plates above them holding the initial column
    (a jittered cubic lattice at 2.2 r), and a closed bin below the outlet
    that catches the discharge.  The hopper's half-width A is sized so the
    column is about twice as tall as it is wide."""
    rng = SplitMix64(seed)
    s = 2.2 * radius
    A = max(0.012, (n * s ** 3 / 16.0) ** (1.0 / 3.0)) + 2.0 * radius + thick
    a = min(0.5 * outlet, 0.5 * A)
    H = A - a                          # 45-degree funnel
    per = max(1, int((2.0 * (A - thick) - 2.0 * radius) / s))
    layers = int(math.ceil(n / (per * per)))
    Hc = layers * s + 4.0 * radius     # column walls above the funnel
    idx = np.arange(n)
    ix, iy, iz = idx % per, (idx // per) % per, idx // (per * per)
    x0 = -(per - 1) * 0.5 * s
    pos = np.stack([x0 + ix * s, x0 + iy * s, H + 2.0 * radius + iz * s], 1)
    pos += rng.uniform(3 * n, -0.05 * radius, 0.05 * radius).reshape(n, 3)
    vel = np.zeros((n, 3))
    sphere = DOModel("Sphere", "Granule", "mobile", "local", density, 1.0, (0, 0, 0), "Sphere", (radius,))
    mass = sphere.analysis()[1]
    Ls = math.hypot(A - a, H)
    d, hh = (A - a) / Ls, H / Ls
    funnel = DOModel("Funnel", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                     (Ls + 2.0 * thick, thick, 2.0 * A + 2.0 * thick))
    column = DOModel("Column", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                     (Hc, thick, 2.0 * A + 2.0 * thick))
    # bin: as wide as the hopper, deep enough for every sphere at a loose packing
    Hb = max(4.0 * radius, 1.6 * n * (2.0 * radius) ** 3 / (2.0 * A) ** 2) + 0.05
    binwall = DOModel("Bin", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                      (2.0 * A + 2.0 * thick, thick, Hb + 2.0 * thick))
    floor = DOModel("Floor", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                    (2.0 * A + 2.0 * thick, thick, 2.0 * A + 2.0 * thick))
    c = 0.5 * (a + A)
    zf = 0.5 * H
    st = [
        # inclined plates: ox up the slope, oz (normal) into the hopper
        _plate("Funnel", (c, 0.0, zf), (d, 0.0, hh), (-hh, 0.0, d)),
        _plate("Funnel", (-c, 0.0, zf), (-d, 0.0, hh), (hh, 0.0, d)),
        _plate("Funnel", (0.0, c, zf), (0.0, d, hh), (0.0, -hh, d)),
        _plate("Funnel", (0.0, -c, zf), (0.0, -d, hh), (0.0, hh, d)),
        # column: vertical plates at |x| = A, |y| = A above the funnel
        _plate("Column", (A, 0.0, H + 0.5 * Hc), (0.0, 0.0, 1.0), (-1.0, 0.0, 0.0)),
        _plate("Column", (-A, 0.0, H + 0.5 * Hc), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)),
        _plate("Column", (0.0, A, H + 0.5 * Hc), (0.0, 0.0, 1.0), (0.0, -1.0, 0.0)),
        _plate("Column", (0.0, -A, H + 0.5 * Hc), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)),
        # bin below the outlet
        _plate("Floor", (0.0, 0.0, -Hb), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)),
        _plate("Bin", (A, 0.0, -0.5 * Hb), (0.0, 1.0, 0.0), (-1.0, 0.0, 0.0)),
        _plate("Bin", (-A, 0.0, -0.5 * Hb), (0.0, 1.0, 0.0), (1.0, 0.0, 0.0)),
        _plate("Bin", (0.0, A, -0.5 * Hb), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0)),
        _plate("Bin", (0.0, -A, -0.5 * Hb), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)),
    ]
    if law == "ISwH":
        pair = hertz_mindlin(E, nu, mu, e, radius, radius, mass, mass)
        wall = hertz_mindlin(E, nu, mu, e, radius, None, mass, None)
        if dt is None:
            v_ref = math.sqrt(2.0 * 9.81 * (H + Hc + Hb))
            k = ish_first_contact_kn(pair["HertzNormalStiffness"], pair["EquivalentMass"],
                                     pair["DampingRatio"], v_ref)
            dt = critical_dt(tangential_mass(pair["EquivalentMass"]), k)
    else:
        pair = spring_dashpot(kn, 0.8, e, mu, True)
        wall = spring_dashpot(kn, 0.8, e, mu, False)
        if dt is None:
            dt = critical_dt(tangential_mass(0.5 * mass), kn)
    iact = [IactModel("Granule", "Granule", law, pair), IactModel("Wall", "Granule", law, wall)]
    world = DOWorld(SystemParameter(0.0, steps * dt, dt, 0.0, (0.0, 0.0, -9.81)),
                    [sphere, funnel, column, binwall, floor], iact, st)
    return _spliced(world, 0, pos, vel, np.zeros((n, 3)), up_iact, record)


def hopper_cfg3(n: int = 4_000_000, seed: int = 6, r_small: float = 0.8e-3, r_large: float = 1.6e-3,
                frac_small: float = 0.7, density: float = 2500.0, E: float = 5e7, nu: float = 0.25,
                mu: float = 0.5, e: float = 0.5, outlet: float = 0.03, dt: float = 2e-6, steps: int = 20000,
                thick: float = 2e-3, zoi_below: float = 2.0, up_iact: int = 20,
                record: int = 10 ** 9) -> WorldArrays:
    """Hopper discharge with a zone of interest: bimodal spheres (r_small /
    r_large, frac_small of them small, by a seeded draw per site), two
    DOModels in two DOGroups with ISwH between every pair of groups and
    against the walls (E, nu, mu, e mapped as in hertz_mindlin); four
    quasi-plates inclined at 45 degrees down to a square outlet of side
    `outlet` at z = 0, four vertical plates above.  The spheres fill the
    funnel and the column on a cubic lattice of 2.2 r_large (jitter
    0.05 r_large), so the lowest fall through the outlet from the start; the
    zone of interest ends zoi_below r_large under the outlet, and spheres
    below it are erased at rebuilds (Consultant::CleanUp).  The hopper's
    half-width A is sized so the filled column is about twice as tall as
    wide."""
    rng = SplitMix64(seed)
    s = 2.2 * r_large
    a = 0.5 * outlet
    A = max(0.03, (n * s ** 3 / 16.0) ** (1.0 / 3.0)) + 2.0 * r_large + thick
    H = A - a                                   # 45-degree funnel
    m_in = (r_large + 0.5 * thick) * math.sqrt(2.0)   # a centre's distance from an inclined mid-plane
    per = int((2.0 * (A - 0.5 * thick - r_large)) / s)
    x0 = -(per - 1) * 0.5 * s
    grid = x0 + s * np.arange(per)
    X, Y = np.meshgrid(grid, grid, indexing="ij")
    X, Y = X.ravel(), Y.ravel()
    sites = []
    count = 0
    z = r_large + 0.5 * s
    while count < n:
        w = (a + z - m_in) if z < H else (A - 0.5 * thick - r_large)
        keep = (np.abs(X) <= w) & (np.abs(Y) <= w)
        k = int(keep.sum())
        if k:
            sites.append(np.stack([X[keep], Y[keep], np.full(k, z)], 1))
            count += k
        z += s
    pos = np.concatenate(sites)[:n]
    pos = pos + rng.uniform(3 * n, -0.05 * r_large, 0.05 * r_large).reshape(n, 3)
    small = rng.uniform(n, 0.0, 1.0) < frac_small
    Hc = float(pos[:, 2].max()) - H + 4.0 * r_large
    m_s = DOModel("Small", "Small", "mobile", "local", density, 1.0, (0, 0, 0), "Sphere", (r_small,))
    m_l = DOModel("Large", "Large", "mobile", "local", density, 1.0, (0, 0, 0), "Sphere", (r_large,))
    mass_s, mass_l = m_s.analysis()[1], m_l.analysis()[1]
    Ls = math.hypot(A - a, H)
    d, hh = (A - a) / Ls, H / Ls
    funnel = DOModel("Funnel", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                     (Ls + 2.0 * thick, thick, 2.0 * A + 2.0 * thick))
    column = DOModel("Column", "Wall", "fixed", "global", 0.0, 1.0, (0, 0, 0), "QuasiPlate",
                     (Hc, thick, 2.0 * A + 2.0 * thick))
    c = 0.5 * (a + A)
    zf = 0.5 * H
    st = [
        _plate("Funnel", (c, 0.0, zf), (d, 0.0, hh), (-hh, 0.0, d)),
        _plate("Funnel", (-c, 0.0, zf), (-d, 0.0, hh), (hh, 0.0, d)),
        _plate("Funnel", (0.0, c, zf), (0.0, d, hh), (0.0, -hh, d)),
        _plate("Funnel", (0.0, -c, zf), (0.0, -d, hh), (0.0, hh, d)),
        _plate("Column", (A, 0.0, H + 0.5 * Hc), (0.0, 0.0, 1.0), (-1.0, 0.0, 0.0)),
        _plate("Column", (-A, 0.0, H + 0.5 * Hc), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)),
        _plate("Column", (0.0, A, H + 0.5 * Hc), (0.0, 0.0, 1.0), (0.0, -1.0, 0.0)),
        _plate("Column", (0.0, -A, H + 0.5 * Hc), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)),
    ]
    iact = [IactModel("Small", "Small", "ISwH", hertz_mindlin(E, nu, mu, e, r_small, r_small, mass_s, mass_s)),
            IactModel("Small", "Large", "ISwH", hertz_mindlin(E, nu, mu, e, r_small, r_large, mass_s, mass_l)),
            IactModel("Large", "Large", "ISwH", hertz_mindlin(E, nu, mu, e, r_large, r_large, mass_l, mass_l)),
            IactModel("Wall", "Small", "ISwH", hertz_mindlin(E, nu, mu, e, r_small, None, mass_s, None)),
            IactModel("Wall", "Large", "ISwH", hertz_mindlin(E, nu, mu, e, r_large, None, mass_l, None))]
    sp = SystemParameter(0.0, steps * dt, dt, 0.0, (0.0, 0.0, -9.81))
    sp.zoi_lower = (0.0, 0.0, -zoi_below * r_large)
    sp.zoi_upper = (0.0, 0.0, H + Hc + 1.0)
    world = DOWorld(sp, [m_s, m_l, funnel, column], iact, st)
    wa = _spliced(world, 0, pos, np.zeros((n, 3)), np.zeros((n, 3)), up_iact, record)
    wa.model_of[:n] = np.where(small, 0, 1).astype(np.int32)
    return wa


def drum_radius_for(n: int, median: float = 1e-3) -> float:
    """Drum radius whose lower half holds n spheres of the median radius at
    a dense packing in a drum as long as its radius."""
    return max(0.02, (2.0 * n * (2.0 * median) ** 3 / (0.7 * math.pi)) ** (1.0 / 3.0))
