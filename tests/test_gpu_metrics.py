"""Long-run statistics on the GPU against the reference and the oracle.

* Wall pressure (vedo_amd.metrics: sum of |ImpactToMaster| over a wall's
  active contacts / (dt x area)) from the GPU's exported records equals, bit
  for bit, the pressure formed from the reference's own ImpactStatus records
  in the golden dumps (spring-dashpot fixtures).
* Ensemble statistics of ISwH runs, which are chaotic (the device pow rounds
  differently from glibc's, and the differences grow): the CPU-reference
  configuration (10,000 spheres on a jittered lattice in a 0.1 m box, Hertz
  E = 1e7 Pa, nu = 0.3, mu = 0.5, e = 0.5, dt = 1e-6 s, 2,000 steps) and a
  lattice of 4,096 spheres pressed together and against all six walls, each
  against the oracle after 2,000 steps: packing fraction (the knight's
  Monte-Carlo CalcDensity, CalcDensity.cpp:11-72, same points) within 0.1
  percentage points, translational and rotational kinetic energy within 1%,
  the kinetic-energy decay from the start within 1%, and the walls' total
  pressure averaged over the second half of the run (every 50 steps: one
  step's pressure on a few contacts is no ensemble quantity) within 5%.
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


def _stats(sim, wa, box):
    st = sim.state()
    et, er = metrics.kinetic_energy(st, wa)
    p = metrics.wall_pressure(*sim.active(), wa, wa.dt)
    return dict(density=sim.packing_density(box, 20000, 999), translation=et, rotation=er, pressure=p)


@pytest.mark.parametrize("kind", ["cpu_reference_config", "dense"])
def test_ish_run_ensemble_statistics_match_oracle(kind):
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
    sph = wa.model_shape[wa.model_of] == 1
    lo, hi = wa.pos[sph].min(0), wa.pos[sph].max(0)
    box = [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]]
    gpu, cpu = GpuSimulation(wa), OracleSim(wa)
    e0 = metrics.kinetic_energy({"vel": wa.vel, "av": wa.av}, wa)
    gpu.step(1000)
    cpu.step(1000)
    pg, pc = [], []
    for _ in range(20):
        gpu.step(50)
        cpu.step(50)
        pg.append(sum(metrics.wall_pressure(*gpu.active(), wa, wa.dt).values()))
        pc.append(sum(metrics.wall_pressure(*cpu.active(), wa, wa.dt).values()))
    g, c = _stats(gpu, wa, box), _stats(cpu, wa, box)
    pmax = max(float(np.mean(pc)), 1e-300)
    dp = abs(float(np.mean(pg)) - float(np.mean(pc)))
    ke0 = e0[0] + e0[1]
    decay_g = (g["translation"] + g["rotation"]) / ke0
    decay_c = (c["translation"] + c["rotation"]) / ke0
    print(f"\n  [{kind}] packing {g['density']:.3f}/{c['density']:.3f} %, translation "
          f"{g['translation']:.6e}/{c['translation']:.6e}, rotation {g['rotation']:.6e}/{c['rotation']:.6e}, "
          f"KE decay {decay_g:.6f}/{decay_c:.6f}, mean total wall pressure {pmax:.4e} Pa (GPU off by {dp:.2e}), "
          f"active contacts {gpu.count_active()}/{cpu.active()[0].size}", flush=True)
    assert abs(g["density"] - c["density"]) <= 0.1
    for f in ("translation", "rotation"):
        assert abs(g[f] - c[f]) <= 1e-2 * abs(c[f]), f
    assert abs(decay_g - decay_c) <= 1e-2 * decay_c
    assert dp <= 0.05 * pmax
    gpu.close()
