"""Wall motion evaluated on the device (vg_set_wall_rotation, a rigid rotation
about a fixed axis) against the host-assigned statuses the reference pins
(SimMediator::Run(const std::map<_uint_t,DOStatus>&), SimMediator.cpp:1121-1163,
fixture drum_kin, which the host-assigned GPU run reproduces bit for bit).

drum_kin's walls are assigned every step from scenes.rigid_rotation_status
(axis through the origin along +y, 20 rad/s, at t = step * dt).  The device
evaluates the same closed form with its own cos/sin and operation order, so
a wall status can differ from the host's in its last bits; from the first
such step on, the two runs are two trajectories of a chaotic system (a
contact law with branches: the Coulomb cap, first contact) and drift apart.
Bars:
  * every step, every wall: position, velocity, Ox, Oz within 1e-15 relative
    (a few ulps) of the host's status;
  * until the first step whose wall statuses differ in any bit, the sphere
    states are bitwise those of the host-assigned run;
  * after 120 steps (about 1.4 ms of drum time): the same number of active
    contacts within 5% and translational / rotational kinetic energy within
    1% (DOWorld::CalculateSystemEnergy on the device).
"""
import numpy as np
import pytest

from conftest import load_golden, step_golden

pytestmark = pytest.mark.gpu

AXIS_POINT, AXIS_DIR, OMEGA = (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 20.0


def _rel(x, ref):
    return float(np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-300))


def test_device_rotation_matches_assigned_walls():
    from vedo_amd.gpu import GpuSimulation
    z, wa = load_golden("drum_kin")
    ids = np.asarray(z["assigned_ids"], np.uint32)
    host = GpuSimulation(wa)
    dev = GpuSimulation(wa)
    dev.set_wall_rotation(ids, AXIS_POINT, AXIS_DIR, OMEGA, 0.0)
    sphere = wa.model_shape[wa.model_of] == 1
    nsteps = int(z["nsteps"])
    first_diff = None
    worst = 0.0
    for step in range(nsteps):
        step_golden(host, z, step)
        dev.step(1)
        a, b = host.state(), dev.state()
        e = max(_rel(b[f][ids], a[f][ids]) for f in ("pos", "vel", "ox", "oz"))
        worst = max(worst, e)
        assert e <= 1e-15, f"step {step + 1}: wall status off by {e:.2e}"
        same_walls = all(np.array_equal(b[f][ids], a[f][ids]) for f in ("pos", "vel", "ox", "oz", "av"))
        if first_diff is None:
            if not same_walls:
                first_diff = step + 1
            else:
                for f in ("pos", "vel", "av"):
                    assert np.array_equal(b[f][sphere], a[f][sphere]), f"step {step + 1}: {f} not bitwise"
    na, nb = host.count_active(), dev.count_active()
    ea, eb = host.system_energy(), dev.system_energy()
    print(f"\n  walls within {worst:.1e} of the host's statuses; first step with a bit difference: {first_diff}; "
          f"after {nsteps} steps active {nb}/{na}, translation {eb['translation']:.6e}/{ea['translation']:.6e}, "
          f"rotation {eb['rotation']:.6e}/{ea['rotation']:.6e}", flush=True)
    assert abs(nb - na) <= 0.05 * na
    for f in ("translation", "rotation"):
        assert abs(eb[f] - ea[f]) <= 1e-2 * abs(ea[f]), f
    host.close()
    dev.close()
