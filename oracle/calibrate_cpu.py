"""CPU calibration of the oracle against the compiled reference.

TEST INFRASTRUCTURE (runs in the build container, where /root/reference and
the reference build oracle/_ref exist).  Times, on one pinned core, the
oracle (oracle/libvedo_oracle.so) and the reference's own SimMediator::Run
(oracle/_ref/vedo_ref_harness, LeapConsultant) on the same bounded sample of
the drum workload that bench.py's cpu_baseline times on the GPU host: the
20,000 spheres of an axial slab in the middle of the 1M-sphere drum bed
(scenes.drum_scene), with the drum walls held fixed.  The harness's set-up
time is removed by timing two step counts and taking the difference.

    python oracle/calibrate_cpu.py [seconds]

writes oracle/cpu_calibration.json.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[0] = ROOT   # the package root, not this folder (oracle.py would shadow the package)

HARNESS = os.path.join(HERE, "_ref", "vedo_ref_harness")


def slab_sample(n=1_000_000, target=20_000, up_iact=100):
    """The bench's CPU sample, taken from the drum's starting bed."""
    from vedo_amd import scenes
    from vedo_amd.world import WorldArrays
    wa, wall_ids, c, a = scenes.drum_scene(n, arrays=True, up_iact=up_iact)
    y = wa.pos[:n, 1]
    keep = np.sort(np.argsort(np.abs(y - float(np.median(y))), kind="stable")[:target])
    ids = np.concatenate([keep, np.arange(n, wa.n_do)])
    sub = WorldArrays(**{k: getattr(wa, k) for k in wa.__dataclass_fields__})
    sub.model_of = wa.model_of[ids]
    sub.model_behavior = np.where(wa.model_shape == 1, wa.model_behavior, 2).astype(np.int32)
    for f in ("pos", "vel", "ox", "oz", "av"):
        setattr(sub, f, np.ascontiguousarray(getattr(wa, f)[ids]))
    return sub, len(keep)


def time_oracle(wa, steps):
    from oracle.oracle import OracleSim
    sim = OracleSim(wa)
    t0 = time.perf_counter()
    sim.step(steps)
    return time.perf_counter() - t0


def time_reference(wa, steps, core):
    from vedo_amd.world import write_vscene
    with tempfile.TemporaryDirectory() as td:
        scene = os.path.join(td, "scene.txt")
        write_vscene(wa, scene, 0)
        cmd = ["taskset", "-c", str(core), HARNESS, scene, os.path.join(td, "out.bin"), str(steps), str(steps)]
        t0 = time.perf_counter()
        subprocess.run(cmd, cwd=td, check=True, stdout=subprocess.DEVNULL)
        return time.perf_counter() - t0


def main():
    seconds = float(sys.argv[1]) if len(sys.argv) > 1 else 15.0
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    core = min(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {core})
    wa, n = slab_sample()
    # step count for about `seconds` of oracle time
    probe = 50
    per = time_oracle(wa, probe) / probe
    steps = max(100, int(seconds / per) // 100 * 100)
    t_or = time_oracle(wa, steps)
    short = steps // 4
    t_ref = time_reference(wa, steps, core) - time_reference(wa, short, core)
    model = None
    for line in open("/proc/cpuinfo"):
        if line.startswith("model name"):
            model = line.split(":", 1)[1].strip()
            break
    out = {"spheres": n, "steps": steps, "core": core, "nproc": os.cpu_count(), "cpu_model": model,
           "oracle_particle_steps_per_s": n * steps / t_or,
           "reference_particle_steps_per_s": n * (steps - short) / t_ref,
           "oracle_vs_reference": (n * steps / t_or) / (n * (steps - short) / t_ref),
           "sample": "20,000 spheres of an axial slab in the middle of the 1M-sphere drum's starting bed, walls "
                     "held fixed, LeapConsultant, UpIact 100; reference time = difference of two harness runs "
                     f"({steps} and {short} steps) on core {core}"}
    with open(os.path.join(HERE, "cpu_calibration.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
