"""Project autotuner for the per-scene tuning knob of the step.

The neighbour-list skin decides how many candidates each sphere walks every
step (more with a larger skin) against how often the lists are rebuilt from
the candidate list (more often with a smaller skin).  Results are identical
for every skin; only the time changes.  ``tune_skin`` times whole steps,
relists and candidate-list rebuilds included, for each candidate skin and
keeps the fastest.
"""
from __future__ import annotations

import time
from typing import Dict, Iterable, Tuple


def tune_skin(sim, r_min: float, up_iact: int, factors: Iterable[float] = (0.1, 0.25, 0.5, 1.0),
              rounds: int = 1) -> Tuple[float, Dict[float, float]]:
    """Time ``rounds * up_iact`` steps per candidate skin (factor * r_min) on
    the live simulation and set the fastest.  Returns (skin, {skin: s/step})."""
    res: Dict[float, float] = {}
    for f in factors:
        skin = f * r_min
        sim.set_skin(skin)
        sim.step(2)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(rounds * up_iact)
        res[skin] = (time.perf_counter() - t0) / (rounds * up_iact)
    best = min(res, key=res.get)
    sim.set_skin(best)
    return best, res
