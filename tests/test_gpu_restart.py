"""IactRecordTab restore on the GPU (vg_set_history, SimMediator::Initiate,
SimMediator.cpp:690-707): restarted at a rebuild step from its DOStatus and
its exported active records (vg_get_active), a run continues bit for bit as
the uninterrupted run and as the oracle's restart from the same records;
the exported records themselves equal the oracle's (flags, shear force, Kn /
initial velocity)."""
import dataclasses

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

CASES = ["box_bsdbf", "box_h", "box_hertz", "box_btsdbf", "box_b2f", "box_piston"]
EXACT = {"box_bsdbf", "box_btsdbf", "box_b2f", "box_piston"}   # spring-dashpot laws: no device pow


def _bits(a, b):
    return int((a.view(np.uint64) != b.view(np.uint64)).sum())


@pytest.mark.parametrize("name", CASES)
def test_gpu_restart_at_rebuild(name):
    from oracle.oracle import OracleSim
    from vedo_amd.gpu import GpuSimulation
    z, wa = load_golden(name)
    k_restart, more = 3 * wa.up_iact, 2 * wa.up_iact + 3
    a = GpuSimulation(wa)
    a.step(k_restart)
    st, t = a.state(), a.time
    lo, hi, rec = a.active()
    print(f"\n[{name}] {len(lo)} records at step {k_restart}", flush=True)
    assert len(lo) > 0
    # the exported records against the oracle's at the same step
    o = OracleSim(wa)
    o.step(k_restart)
    olo, ohi, orec = o.active()
    order = np.lexsort((ohi, olo))   # the oracle lists them in interaction order
    olo, ohi, orec = olo[order], ohi[order], orec[order]
    assert np.array_equal(lo, olo) and np.array_equal(hi, ohi)
    for f in ("contact", "last_contact", "bond", "last_bond"):
        assert np.array_equal(rec[f], orec[f]), f
    if name in EXACT:
        assert _bits(rec["shear"], orec["shear"]) == 0
    else:
        assert np.allclose(rec["shear"], orec["shear"], rtol=1e-9, atol=1e-300)
    w2 = dataclasses.replace(wa, pos=st["pos"], vel=st["vel"], ox=st["ox"], oz=st["oz"], av=st["av"], t_current=t)
    b = GpuSimulation(w2)
    b.set_history(lo, hi, rec)
    ob = OracleSim(w2)
    ob.set_history(lo, hi, rec)
    a.step(more)
    b.step(more)
    ob.step(more)
    sa, sb, so = a.state(), b.state(), ob.state()
    for f in ("pos", "vel", "av"):
        assert _bits(sa[f], sb[f]) == 0, f"{f}: restart differs from the uninterrupted run"
        if name in EXACT:
            assert _bits(sb[f], so[f]) == 0, f"{f}: GPU restart differs from the oracle's"
        else:
            assert np.abs(sb[f] - so[f]).max() <= 1e-9 * max(np.abs(so[f]).max(), 1e-300)
    a.close()
    b.close()


def test_gpu_set_history_is_validated():
    from vedo_amd.gpu import GpuError, GpuSimulation
    z, wa = load_golden("box_bsdbf")
    sim = GpuSimulation(wa)
    lo, hi, rec = sim.active()
    with pytest.raises(GpuError):   # a DO paired with itself
        sim.set_history(np.array([3], np.uint32), np.array([3], np.uint32), rec[:1] if len(rec) else
                        np.zeros(1, rec.dtype))
    sim.step(1)
    with pytest.raises(GpuError):   # only before the first step
        sim.set_history(lo, hi, rec)
    sim.close()
