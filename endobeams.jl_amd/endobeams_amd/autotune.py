"""Autotuning of the batch shape and the launch shapes of a design sweep.

A sweep of independent structures can be advanced as one batch context or as several smaller
ones. A larger batch fills the GPU better; a smaller one iterates only until its own slowest
instance has converged (each instance keeps its own Δt either way). `pick_batch_size` measures
instead of guessing: for each candidate size it builds a context of that many instances (the
first ones of the sweep), runs warm-up steps, lets the library's own autotuner (eb_autotune) pick
the launch shapes on the live state, and times converged steps; the candidate with the highest
element-steps/s wins.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from . import abi
from .engine import HipBatchEngine


def time_batch(instances, precision=abi.EB_PRECISION_FP32, warmup=1, steps=2, tune=True):
    """Element-steps/s of one context holding `instances` (each instance with its own Δt,
    eb_run_steps_batch), with its launch shapes."""
    eng = HipBatchEngine(instances, precision=precision)
    try:
        eng.initialize_state_simulation()
        states = eng.run_steps_batch(warmup)
        shapes = eng.autotune()[0] if tune else eng.tuning
        E = [int(eng.beam_off[s + 1] - eng.beam_off[s]) for s in range(eng.S)]
        n0 = [r.steps for r in states]
        t0 = time.perf_counter()
        states = eng.run_steps_batch(steps, states)
        wall = time.perf_counter() - t0
        return sum(e * (r.steps - n) for e, r, n in zip(E, states, n0)) / wall, shapes
    finally:
        eng.close()


def pick_batch_size(instances: Sequence, candidates: Sequence[int], precision=abi.EB_PRECISION_FP32,
                    warmup=1, steps=2, log: Optional[Callable[[int, float], None]] = None
                    ) -> Tuple[int, Dict[int, float], Dict[int, dict]]:
    """Returns (best size, {size: element-steps/s}, {size: launch shapes}) over the candidates
    that do not exceed the number of instances; `log(size, rate)` is called after each."""
    rates, shapes = {}, {}
    sizes = sorted({int(c) for c in candidates if 0 < int(c) <= len(instances)}) or [len(instances)]
    for c in sizes:
        rates[c], shapes[c] = time_batch(list(instances[:c]), precision, warmup, steps)
        if log is not None:
            log(c, rates[c])
    best = max(rates, key=rates.get)
    return best, rates, shapes


def split(instances: Sequence, size: int) -> List[list]:
    """Consecutive groups of `size` instances (the last one may be smaller)."""
    return [list(instances[i:i + size]) for i in range(0, len(instances), size)]
