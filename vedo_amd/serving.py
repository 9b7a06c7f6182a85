"""Continuous batching of DEM scenes on one GPU.

Requests (a scene and a step count) arrive over time.  The server keeps one
batch handle (vg_batch_create) whose device layout holds every running
scene; it admits queued requests whenever their HBM footprint fits the free
capacity, steps all running scenes together one rebuild period at a time
(scenes are placed and removed only at rebuild boundaries), and hands back a
scene's final DOStatus as soon as it has left the device.  Each result is the
same as running the scene alone (see tests/test_gpu_batch.py).

The request stream of the serving configuration: Poisson arrivals, scene kind
drum / hopper / box, sphere count log-uniform, step count uniform, law ISwH or
ISwBSDBF.  Streams are generated from the project's SplitMix64, so a seed
fixes them on every host.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import scenes
from .scenes import SplitMix64

SCENE_DONE = 5     # vg_scene_status: the result is on the host (VG_SCENE_DONE)
SCENE_FAILED = 6   # the scene blew up and left the batch (VG_SCENE_FAILED)


# Device bytes one sphere slot of a batch takes (two state buffers, two
# neighbour lists of `ki` entries with history, candidate list, rebuild
# scratch); the footprint of a scene is n * slot_bytes.
SKIN_FLOOR, SKIN_CAP = 1e-4, 1e-3   # neighbour-list skin bounds of the served scenes (m)


def slot_bytes(ki: int = 32, ko: int = 160) -> int:
    state = 2 * (3 * 32 + 48 + 8)
    lists = 2 * (1 + 2 * ki * 4 + 2 * ki * 32)
    return state + lists + 4 * ko + 8 * 16


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the stream starts
    kind: str               # "drum" | "hopper" | "box"
    n: int
    steps: int
    law: str
    seed: int
    wa: object = None       # WorldArrays, built by build_scene
    walls: Optional[np.ndarray] = None   # kinematic walls (drum), DO ids
    # filled by the server
    t_admit: float = float("nan")
    t_done: float = float("nan")
    failed: bool = False
    result: Optional[Dict[str, np.ndarray]] = None

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


def poisson_requests(count: int, rate: float = 20.0, seed: int = 5, n_range=(5e3, 5e5), steps_range=(2000, 20000),
                     kinds=("drum", "hopper", "box"), laws=("ISwH", "ISwBSDBF")) -> List[Request]:
    """`count` requests with exponential inter-arrival times of mean 1/rate,
    sphere counts log-uniform in n_range, step counts uniform in steps_range."""
    rng = SplitMix64(seed)
    u = rng.uniform(5 * count, 0.0, 1.0).reshape(count, 5)
    t = 0.0
    out = []
    for k in range(count):
        t += -math.log(1.0 - u[k, 0]) / rate
        n = int(round(math.exp(math.log(n_range[0]) + u[k, 1] * (math.log(n_range[1]) - math.log(n_range[0])))))
        steps = int(steps_range[0] + u[k, 2] * (steps_range[1] - steps_range[0]))
        kind = kinds[min(int(u[k, 3] * len(kinds)), len(kinds) - 1)]
        law = laws[min(int(u[k, 4] * len(laws)), len(laws) - 1)]
        out.append(Request(k, t, kind, n, steps, law, seed * 1000 + k))
    return out


def build_scene(req: Request, up_iact: int = 20) -> Request:
    """The request's WorldArrays: drums rotate at 2 rad/s (kinematic walls),
    hoppers discharge, boxes settle; materials of the configurations."""
    if req.kind == "drum":
        R = scenes.drum_radius_for(req.n)
        if req.law == "ISwH":
            wa, walls, c, a = scenes.drum_scene(req.n, seed=req.seed, R=R, arrays=True, up_iact=up_iact,
                                                steps=req.steps, law="ISwH")
        else:
            wa, walls, c, a = scenes.drum_scene(req.n, seed=req.seed, R=R, arrays=True, up_iact=up_iact,
                                                steps=req.steps)
        req.walls = np.asarray(walls, np.uint32)
    elif req.kind == "hopper":
        wa = scenes.hopper_arrays(req.n, seed=req.seed, law=req.law, steps=req.steps, up_iact=up_iact)
    else:
        rng = SplitMix64(req.seed)
        mu, e, lE = rng.uniform(3, 0.0, 1.0)
        wa = scenes.box_arrays(req.n, seed=req.seed, law=req.law, mu=0.2 + 0.5 * mu, e=0.3 + 0.6 * e,
                               E=10.0 ** (6.0 + 2.0 * lE), steps=req.steps, up_iact=up_iact)
    req.wa = wa
    return req


@dataclass
class ServeStats:
    requests: int
    wall: float
    particle_steps: float
    latencies: List[float] = field(default_factory=list)
    failed: int = 0

    @property
    def throughput(self) -> float:
        return self.particle_steps / self.wall

    def pct(self, q: float) -> float:
        return float(np.percentile(self.latencies, q)) if self.latencies else float("nan")


class Server:
    """Continuous-batching loop over one batch handle.

    capacity: sphere slots of the device layout (its HBM footprint is
    capacity * slot_bytes(); pick it from the memory budget)."""

    def __init__(self, capacity: int, up_iact: int = 20, max_walls: int = 8192, max_scenes: int = 1024,
                 neighbour_capacity: int = 32, keep_results: bool = True, batch=None):
        if batch is None:
            from .gpu import GpuBatch
            batch = GpuBatch(capacity, max_walls=max_walls, max_scenes=max_scenes, up_iact=up_iact,
                             neighbour_capacity=neighbour_capacity)
        self.batch = batch
        # each scene's neighbour-list skin follows its speed (0.1 to 1 mm for
        # these millimetre spheres): few relists between rebuilds
        self.batch.set_skin_auto(SKIN_FLOOR, SKIN_CAP)
        self.capacity = capacity
        self.up_iact = up_iact
        self.keep_results = keep_results

    def close(self):
        self.batch.close()

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            log: Optional[Callable[[str], None]] = None) -> ServeStats:
        pend = sorted(requests, key=lambda r: r.arrival)
        queue: List[Request] = []
        running: Dict[int, Request] = {}
        used = 0
        t0 = clock()
        ps = 0.0
        periods = 0
        while pend or queue or running:
            now = clock() - t0
            while pend and pend[0].arrival <= now:
                queue.append(pend.pop(0))
            # admission by footprint, first come first served
            while queue and used + queue[0].n <= self.capacity:
                r = queue.pop(0)
                sid = self.batch.admit(r.wa, r.steps)
                if r.walls is not None:
                    self.batch.set_wall_rotation(sid, r.walls, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 2.0, r.wa.t_current)
                r.t_admit = clock() - t0
                running[sid] = r
                used += r.n
            if queue and not running:
                raise RuntimeError(f"request {queue[0].rid} ({queue[0].n} spheres) exceeds the capacity")
            if not running:
                # idle until the next arrival
                wait = pend[0].arrival - (clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                continue
            self.batch.step(self.up_iact)
            periods += 1
            for sid in list(running):
                state, done, _ = self.batch.status(sid)
                if state in (SCENE_DONE, SCENE_FAILED):
                    r = running.pop(sid)
                    r.failed = state == SCENE_FAILED
                    if self.keep_results and not r.failed:
                        r.result = self.batch.state(sid)
                    r.t_done = clock() - t0
                    self.batch.retire(sid)
                    used -= r.n
                    if not r.failed:
                        ps += float(r.n) * r.steps
                    if log:
                        log(f"request {r.rid} ({r.kind}, {r.law}, {r.n} spheres, {r.steps} steps) "
                            f"{'FAILED' if r.failed else 'done'} {r.latency:.2f} s after arrival")
        wall = clock() - t0
        ok = [r for r in requests if not r.failed]
        return ServeStats(len(requests), wall, ps, [r.latency for r in ok], failed=len(requests) - len(ok))


def serve_one_by_one(requests: List[Request], clock: Callable[[], float] = time.perf_counter) -> ServeStats:
    """The plain per-scene GPU loop: each request alone, in arrival order,
    one scene on the GPU at a time (GpuSimulation per scene)."""
    from .gpu import GpuError, GpuSimulation
    t0 = clock()
    ps = 0.0
    failed = 0
    lat = []
    for r in sorted(requests, key=lambda q: q.arrival):
        wait = r.arrival - (clock() - t0)
        if wait > 0:
            time.sleep(wait)
        sim = GpuSimulation(r.wa)
        sim.set_skin_auto(SKIN_FLOOR, SKIN_CAP)   # the server's skin policy
        try:
            if r.walls is not None:
                sim.set_wall_rotation(r.walls, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 2.0, r.wa.t_current)
            sim.step(r.steps)
            r.result = sim.state()
            ok = True
        except GpuError:   # the scene blew up (candidate-list overflow)
            ok = False
            failed += 1
        sim.close()
        r.t_done = clock() - t0
        if ok:
            ps += float(r.n) * r.steps
            lat.append(r.latency)
    wall = clock() - t0
    return ServeStats(len(requests), wall, ps, lat, failed=failed)
