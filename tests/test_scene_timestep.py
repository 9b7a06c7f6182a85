"""The batched configuration's time-step rule against the physics it must keep
stable, on the oracle (the FP64 restatement the golden fixtures pin to the
compiled reference).

Scene 101 of the batched configuration (1,298 spheres, E = 1.31e6 Pa,
e = 0.84: soft and weakly damped) is the smallest of the scenes that blew up
with a time step of 0.1 x 2 pi sqrt(M*/k) (M* the pair's reduced mass, k the
ISwH first-contact stiffness at the fall speed): the compiled reference
itself gains energy on it until a sphere crosses more than its radius in one
step (step 640 of 1000).  The contact's tangential-rotational mode is the
stiffer one (effective mass 2/7 M*, scenes.tangential_mass); with its time
step the same scene settles over the same simulated time."""
import numpy as np

from oracle.oracle import OracleSim
from vedo_amd import scenes

RADIUS = 1e-3


def _max_speed_steps(wa, n, steps, every=20):
    sim = OracleSim(wa)
    sph = wa.model_shape[wa.model_of] == 1
    out = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        sim.step(k)
        done += k
        v = np.linalg.norm(sim.state()["vel"][sph], axis=1).max()
        out.append((done, float(v)))
        if v * wa.dt > RADIUS:
            break
    return out


def test_normal_mode_time_step_blows_up_and_tangential_one_settles():
    wa = scenes.batched_box(101, 1000)
    n = int((wa.model_shape[wa.model_of] == 1).sum())
    assert n == 1298
    # the normal mode's time step (7/2 the mass: sqrt(3.5) longer), 1000 steps
    old = scenes.batched_box(101, 1000)
    old.dt = wa.dt * np.sqrt(3.5)
    old.t_stop = old.t_start + 1000 * old.dt
    hist_old = _max_speed_steps(old, n, 1000)
    print(f"\n  normal-mode dt {old.dt:.3e}: blown up at step {hist_old[-1][0]}, v_max {hist_old[-1][1]:.2f} m/s")
    assert hist_old[-1][1] * old.dt > RADIUS, "the normal-mode time step no longer blows this scene up"
    # the rule's time step over the same simulated time
    steps = int(np.ceil(1000 * old.dt / wa.dt))
    wa.t_stop = wa.t_start + steps * wa.dt
    hist = _max_speed_steps(wa, n, steps)
    vmax = max(v for _, v in hist)
    print(f"  tangential-mode dt {wa.dt:.3e}: {hist[-1][0]} steps, largest v_max {vmax:.2f} m/s")
    assert hist[-1][0] == steps
    assert vmax < 1.0   # a fall through the lattice and the initial velocities stay below 1 m/s
