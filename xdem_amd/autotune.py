"""Autotuner of the engine's two free sizes, timed on the workload itself.

* skin: the pair-list skin, which also sets the list grid's cell
  (2 r_max + skin).  A small skin lists fewer pairs per step but rebuilds
  the list more often; the best value depends on the packing's mobility.
* pair_blocks: workgroups of the pair kernels (coarse pass, contact,
  guess), i.e. how many listed pairs a workgroup serves per sweep.

Neither changes any result (contact sets, forces and their order are fixed
by the reference's rules), so the tuner only times.  Candidates are scored
by the wall time of `steps` steps after `warmup` steps, from the same state.
"""
from __future__ import annotations

import itertools
import time

from .engine import Engine


def autotune(states, precision: int = 0, skin_factors=(0.1, 0.2, 0.35),
             blocks_per_cu=(4, 8, 16), cus: int = 256, steps: int = 40, warmup: int = 10,
             log=None):
    """Returns ({"skin_factor", "pair_blocks", "ms_per_step"}, all trials).
    skin = skin_factor * (largest outer radius of each packing)."""
    import numpy as np
    trials = []
    for sf, bpc in itertools.product(skin_factors, blocks_per_cu):
        trial = []
        for s in states:
            t = s.copy()
            t.cfg.skin = sf * float(np.max(np.abs(s.r)))
            t.cfg.pair_blocks = bpc * cus
            trial.append(t)
        eng = Engine(trial, precision=precision)
        eng.step(warmup)
        eng.sync()
        t0 = time.perf_counter()
        eng.step(steps, sync=False)
        eng.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        eng.close()
        trials.append({"skin_factor": sf, "pair_blocks": bpc * cus, "ms_per_step": ms})
        if log:
            log(f"[autotune] skin {sf:.2f} r_max, {bpc * cus} blocks: {ms:.3f} ms/step")
    best = min(trials, key=lambda d: d["ms_per_step"])
    return best, trials
