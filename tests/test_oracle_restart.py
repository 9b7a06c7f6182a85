"""IactRecordTab restore (SimMediator::Initiate, SimMediator.cpp:690-707) in the
oracle: a run restarted at a rebuild step from its DOStatus and its active
records continues bit for bit as the uninterrupted run, whose rebuild at that
step carries the histories through the same IactRecordTab
(NBSConsultant.cpp:376-395); the uninterrupted run is pinned to the compiled
reference by the golden fixtures.  This pins vo_set_history."""
import dataclasses

import numpy as np
import pytest

from conftest import load_golden

CASES = ["box_bsdbf", "box_h", "box_hertz", "box_btsdbf", "box_b2f", "box_piston"]


def _restart_world(wa, st, t):
    return dataclasses.replace(wa, pos=st["pos"], vel=st["vel"], ox=st["ox"], oz=st["oz"], av=st["av"],
                               t_current=t)


@pytest.mark.parametrize("name", CASES)
def test_oracle_restart_at_rebuild_is_transparent(name):
    from oracle.oracle import OracleSim
    z, wa = load_golden(name)
    k_restart = 3 * wa.up_iact
    more = 2 * wa.up_iact + 3
    a = OracleSim(wa)
    a.step(k_restart)
    st, t = a.state(), a.time
    lo, hi, rec = a.active()
    assert len(lo) > 0
    b = OracleSim(_restart_world(wa, st, t))
    b.set_history(lo, hi, rec)
    a.step(more)
    b.step(more)
    sa, sb = a.state(), b.state()
    for f in ("pos", "vel", "ox", "oz", "av"):
        assert np.array_equal(sa[f].view(np.uint64), sb[f].view(np.uint64)), f


def test_oracle_restart_without_records_differs():
    # the records matter: dropping them (all statuses ImpactStatus()) changes the run
    from oracle.oracle import OracleSim
    z, wa = load_golden("box_bsdbf")
    a = OracleSim(wa)
    a.step(3 * wa.up_iact)
    st, t = a.state(), a.time
    lo, hi, rec = a.active()
    assert (rec["shear"] != 0).any()
    b = OracleSim(_restart_world(wa, st, t))
    a.step(10)
    b.step(10)
    assert not np.array_equal(a.state()["vel"], b.state()["vel"])
