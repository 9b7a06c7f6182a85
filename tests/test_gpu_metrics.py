"""Long-run statistics on the GPU against the reference and the oracle.

* Wall pressure (vedo_amd.metrics: sum of |ImpactToMaster| over a wall's
  active contacts / (dt x area)) from the GPU's exported records equals, bit
  for bit, the pressure formed from the reference's own ImpactStatus records
  in the golden dumps (spring-dashpot fixtures).
* Ensemble statistics of ISwH runs, which are chaotic (the device's pow and
  cbrt round differently from glibc's pow, and the differences grow): the
  CPU-reference configuration (10,000 spheres on a jittered lattice in a
  0.1 m box, Hertz E = 1e7 Pa, nu = 0.3, mu = 0.5, e = 0.5, dt = 1e-6 s,
  2,000 steps) and a lattice of 4,096 spheres pressed together and against
  all six walls, each against the oracle after 2,000 steps: packing fraction
  (the knight's Monte-Carlo CalcDensity, CalcDensity.cpp:11-72, same points)
  within 0.1 percentage points; translational and rotational kinetic energy,
  the kinetic-energy decay from the start and the walls' total pressure
  averaged over the second half of the run (every 50 steps: one step's
  pressure on a few contacts is no ensemble quantity) within twice the
  statistic's own spread, or 1% (kinetic energies and decay) and 5%
  (pressure) where the spread is smaller.  The spread is measured on the
  oracle: three more oracle runs start from the same state with one sphere's
  x velocity moved by one ulp, and the spread is the largest difference of
  their statistic from the unperturbed run's.  (On the pressed lattice a
  one-ulp start moves the oracle's translational energy by up to 1.6% and
  its rotational energy by up to 2.9% after 2,000 steps.)
"""
import numpy as np
import pytest

from conftest import load_golden, step_golden
from vedo_amd import metrics, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["box_bsdbf", "box_cyl", "box_orbital", "box_piston", "box_hole", "drum_kin"])
def test_gpu_wall_pressure_equals_reference_records(name):
    from vedo_amd.gpu import GpuSimulation
    z, wa = load_golden(name)
    sim = GpuSimulation(wa)
    step, checked = 0, 0
    for k, d in enumerate(z["dumps"]):
        while step < d:
            step_golden(sim, z, step)
            step += 1
        if d == 0:
            continue
        ri = z[f"d{k}_iact"]
        act = ri["active"].astype(bool)
        ref = metrics.wall_pressure(np.minimum(ri["slave"], ri["master"])[act],
                                    np.maximum(ri["slave"], ri["master"])[act], z[f"d{k}_status"][act], wa, wa.dt)
        mine = metrics.wall_pressure(*sim.active(), wa, wa.dt)
        for w in ref:
            assert np.float64(mine[w]).view(np.uint64) == np.float64(ref[w]).view(np.uint64), (name, d, w)
        checked += 1
    print(f"\n  [{name}] wall pressure bit-exact at {checked} dumps", flush=True)
    sim.close()


def _run(sim, wa, box):
    """The statistics of one run of 2,000 steps (the pressure sampled every
    50 steps over the second half)."""
    e0 = metrics.kinetic_energy({"vel": wa.vel, "av": wa.av}, wa)
    sim.step(1000)
    ps = []
    for _ in range(20):
        sim.step(50)
        ps.append(sum(metrics.wall_pressure(*sim.active(), wa, wa.dt).values()))
    et, er = metrics.kinetic_energy(sim.state(), wa)
    return dict(density=sim.packing_density(box, 20000, 999), translation=et, rotation=er,
                decay=(et + er) / (e0[0] + e0[1]), pressure=float(np.mean(ps)))


@pytest.mark.parametrize("kind", ["cpu_reference_config", "dense"])
def test_ish_run_ensemble_statistics_match_oracle(kind):
    import copy
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import OracleSim
    from vedo_amd.gpu import GpuSimulation
    if kind == "cpu_reference_config":
        w = scenes.box_scene(10000, seed=1, law="ISwH", dt=1e-6, steps=2000)
    else:
        # a lattice a little compressed against all six walls: every contact
        # and every wall loaded from the start, the kinetic energy decays
        w = scenes.box_scene(4096, seed=2, law="ISwH", dt=1e-6, steps=2000, spacing=1.995, jitter=0.01,
                             vmax=0.05, offset=1.995, side=0.03)
    wa = w.to_arrays(up_iact=20, record=10 ** 9)
    sph = np.nonzero(wa.model_shape[wa.model_of] == 1)[0]
    lo, hi = wa.pos[sph].min(0), wa.pos[sph].max(0)
    box = [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]]

    def oracle(bump):
        w2 = copy.deepcopy(wa)
        if bump is not None:
            i = sph[bump]
            w2.vel[i, 0] = np.nextafter(w2.vel[i, 0], np.inf)
        return _run(OracleSim(w2), w2, box)

    # the oracle runs in threads (its steps release the GIL) beside the GPU's
    with ThreadPoolExecutor(4) as ex:
        fut = [ex.submit(oracle, b) for b in (None, 0, len(sph) // 2, len(sph) - 1)]
        gpu = GpuSimulation(wa)
        g = _run(gpu, wa, box)
        runs = [f.result() for f in fut]
    c = runs[0]
    print(f"\n  [{kind}] GPU / oracle (largest one-ulp spread): packing {g['density']:.3f} / {c['density']:.3f} %, "
          + ", ".join(f"{f} {g[f]:.6e} / {c[f]:.6e} ({max(abs(r[f] - c[f]) for r in runs[1:]):.2e})"
                      for f in ("translation", "rotation", "decay", "pressure"))
          + f", active contacts {gpu.count_active()}", flush=True)
    assert abs(g["density"] - c["density"]) <= 0.1
    for f, floor in (("translation", 1e-2), ("rotation", 1e-2), ("decay", 1e-2), ("pressure", 5e-2)):
        spread = max(abs(r[f] - c[f]) for r in runs[1:])
        assert abs(g[f] - c[f]) <= max(floor * abs(c[f]), 2.0 * spread), f
    gpu.close()
