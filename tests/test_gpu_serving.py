"""Serving (continuous batching): requests of mixed kind, law and size arrive
while others run; each result equals the same scene run alone on the GPU
(bitwise), whatever else shared the device with it."""
import numpy as np
import pytest

from vedo_amd import serving

pytestmark = pytest.mark.gpu


def test_served_results_equal_each_scene_alone():
    from vedo_amd.gpu import GpuSimulation
    reqs = serving.poisson_requests(6, rate=300.0, seed=11, n_range=(300, 3000), steps_range=(40, 160))
    for k, r in enumerate(reqs):   # every kind and law
        r.kind = ("drum", "hopper", "box")[k % 3]
        r.law = ("ISwH", "ISwBSDBF")[k % 2]
        serving.build_scene(r, up_iact=10)
    # capacity below the total: later requests wait for earlier ones to leave
    cap = max(r.n for r in reqs) + 1000
    print(f"\n  {[(r.kind, r.law, r.n, r.steps) for r in reqs]}, capacity {cap}", flush=True)
    srv = serving.Server(capacity=cap, up_iact=10, max_walls=512, max_scenes=16)
    st = srv.run(reqs, log=lambda m: print("  " + m, flush=True))
    srv.close()
    assert len(st.latencies) == len(reqs) and all(np.isfinite(st.latencies))
    for r in reqs:
        sim = GpuSimulation(r.wa)
        if r.walls is not None:
            sim.set_wall_rotation(r.walls, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 2.0, r.wa.t_current)
        sim.step(r.steps)
        ref = sim.state()
        sim.close()
        sph = r.wa.model_shape[r.wa.model_of] == 1
        same = all(np.array_equal(r.result[f][sph].view(np.uint64), ref[f][sph].view(np.uint64))
                   for f in ("pos", "vel", "ox", "oz", "av"))
        print(f"  request {r.rid} ({r.kind}, {r.law}, n={r.n}, {r.steps} steps): bitwise equal to alone: {same}",
              flush=True)
        assert same
