"""vg_scene_packing_density_noalloc (no device allocation, a fixed per-GPU
scratch array) against vg_scene_packing_density, through the C-ABI.

The two entry points share the hit-test kernel and draw the same points, so
their results agree bit for bit: two calls in a row, the second with fewer
points over a smaller box (the scratch keeps the first call's points and
counter; the second must not see them), a call at the scratch's maximum, and
a call above it, which is refused with an error and leaves the next call
working."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

MAX_POINTS = 65536   # VG_DENSITY_NOALLOC_MAX_POINTS


def _call(L, h, fn, box, n, seed=999):
    b = np.ascontiguousarray(box, np.float64)
    out = C.c_double(-1.0)
    rc = getattr(L, fn)(h, 0, b.ctypes.data_as(C.POINTER(C.c_double)), int(n), int(seed), C.byref(out))
    return rc, out.value


def test_packing_density_noalloc_matches_allocating_entry_point():
    from vedo_amd.gpu import GpuSimulation
    z, wa = load_golden("box_bsdbf")
    sim = GpuSimulation(wa)
    sim.step(50)
    L, h = sim.L, sim.h
    p = sim.state()["pos"][wa.model_shape[wa.model_of] == 1]
    lo, hi = p.min(0), p.max(0)
    whole = np.array([lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]])
    lower = whole.copy()
    lower[5] = 0.5 * (lo[2] + hi[2])
    for box, n, seed in ((whole, 20000, 999), (lower, 3000, 7), (whole, MAX_POINTS, 999)):
        rc_a, ref = _call(L, h, "vg_scene_packing_density", box, n, seed)
        rc_b, got = _call(L, h, "vg_scene_packing_density_noalloc", box, n, seed)
        print(f"\n  {n} points: allocating {ref!r}, no-alloc {got!r}", flush=True)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b)
        assert 0.0 < ref < 100.0
        assert np.float64(got).view(np.uint64) == np.float64(ref).view(np.uint64)
    rc, _ = _call(L, h, "vg_scene_packing_density_noalloc", whole, MAX_POINTS + 1)
    assert rc == -1
    assert b"at most" in L.vg_last_error(h)
    rc_a, ref = _call(L, h, "vg_scene_packing_density", lower, 5000, 3)
    rc_b, got = _call(L, h, "vg_scene_packing_density_noalloc", lower, 5000, 3)
    assert rc_a == 0 and rc_b == 0 and got == ref
    sim.close()
