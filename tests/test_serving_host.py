"""Host logic of the serving loop (no GPU): the request stream, footprint
admission and the bookkeeping of results, driven by a stand-in batch that
follows the batch handle's rules (placement at rebuild boundaries, a scene
leaves the device at the boundary after it finished)."""
import math

import numpy as np

from vedo_amd import serving


def test_poisson_stream_is_deterministic_and_in_range():
    a = serving.poisson_requests(400, rate=20.0, seed=9)
    b = serving.poisson_requests(400, rate=20.0, seed=9)
    assert [(r.arrival, r.n, r.steps, r.kind, r.law) for r in a] == [(r.arrival, r.n, r.steps, r.kind, r.law) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1.0 / 20.0) < 0.15 / 20.0          # exponential, mean 1/rate
    assert abs(np.median(gaps) - math.log(2) / 20.0) < 0.25 / 20.0
    n = np.array([r.n for r in a])
    assert n.min() >= 5e3 and n.max() <= 5e5
    ln = np.log(n)
    assert abs(np.median(ln) - 0.5 * (math.log(5e3) + math.log(5e5))) < 0.4   # log-uniform
    steps = np.array([r.steps for r in a])
    assert steps.min() >= 2000 and steps.max() <= 20000
    assert {r.kind for r in a} == {"drum", "hopper", "box"} and {r.law for r in a} == {"ISwH", "ISwBSDBF"}


class FakeBatch:
    """Stand-in for GpuBatch: scenes queue, are placed at the next boundary,
    finish after `steps` steps and are DONE at the boundary after that."""

    def __init__(self, up_iact):
        self.U = up_iact
        self.step_count = 0
        self.sc = {}
        self.next = 0
        self.admitted_at = []

    def admit(self, wa, steps):
        sid = self.next
        self.next += 1
        self.sc[sid] = dict(state=1, steps=steps, done=0, n=wa.n)
        self.admitted_at.append(self.step_count)
        return sid

    def set_wall_rotation(self, *a):
        pass

    def set_skin_auto(self, floor, cap):
        assert 0 < floor <= cap

    def step(self, k):
        for _ in range(k):
            for s in self.sc.values():
                if s["state"] == 2:
                    s["done"] += 1
                    if s["done"] == s["steps"]:
                        s["state"] = 3
            self.step_count += 1
            if self.step_count % self.U == 0:
                for s in self.sc.values():
                    if s["state"] == 3:
                        s["state"] = 5
                    elif s["state"] == 1:
                        s["state"] = 2

    def status(self, sid):
        s = self.sc[sid]
        return s["state"], s["done"], 0.0

    def state(self, sid):
        return {"pos": np.zeros((1, 3))}

    def retire(self, sid):
        del self.sc[sid]

    def close(self):
        pass


class WA:
    def __init__(self, n):
        self.n = n
        self.t_current = 0.0


def test_server_admits_by_footprint_and_returns_every_result():
    reqs = serving.poisson_requests(30, rate=50.0, seed=3, n_range=(10, 100), steps_range=(5, 40))
    for r in reqs:
        r.wa = WA(r.n)
        r.walls = None
    clock = [0.0]

    def tick():
        clock[0] += 0.001
        return clock[0]

    fb = FakeBatch(4)
    srv = serving.Server(capacity=150, up_iact=4, batch=fb)
    st = srv.run(reqs, clock=tick)
    assert st.requests == 30 and len(st.latencies) == 30
    assert all(r.result is not None and r.t_done >= r.t_admit >= r.arrival for r in reqs)
    assert st.particle_steps == sum(r.n * r.steps for r in reqs)
    # scenes are placed only at boundaries; each finishes its own steps
    assert all(s % 1 == 0 for s in fb.admitted_at)
    assert not fb.sc   # every scene retired
    assert st.pct(50) <= st.pct(99)
