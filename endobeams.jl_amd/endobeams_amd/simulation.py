"""run_simulation — the host time loop of src/run_simulation.jl:1-119, driving an Engine.

The loop (adaptive Δt, rollback on Newton failure, abort below min_timestep, wall-clock stop) is
the reference's caller of the hot path; it stays on the host and calls the engine once per step.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from .engine import Engine, HipEngine
from .model import BeamsConfiguration, RigidInteraction, SimulationParams


class SimulationAborted(RuntimeError):
    pass


@dataclass
class RunStats:
    steps: int = 0                 # converged steps
    failed_steps: int = 0          # steps retried with Δt/2
    newton_iterations: int = 0     # Σ (k-1) over converged steps
    t_end: float = 0.0
    wall_s: float = 0.0
    iterations_per_step: List[int] = field(default_factory=list)

    @property
    def newton_per_step(self) -> float:
        return self.newton_iterations / max(self.steps, 1)


def run_simulation(conf: BeamsConfiguration, params: SimulationParams, inter: Optional[RigidInteraction] = None,
                   engine: Optional[Engine] = None, on_converged: Optional[Callable] = None,
                   max_steps: Optional[int] = None) -> RunStats:
    """run_simulation!(conf, params, inter) (src/run_simulation.jl:1-119).

    `engine` defaults to the HIP engine. `on_converged(step, t, engine)` is called after each
    converged step, before update_converged! (where the reference writes VTK output).
    `max_steps` bounds the number of converged steps (None: run to simulation_end_time)."""
    eng = engine if engine is not None else HipEngine(conf, params, inter)
    stats = RunStats()
    tn = 0.0
    dt = params.initial_timestep
    successive = 0
    step = 1
    start = time.perf_counter()
    eng.initialize_state_simulation()
    while params.simulation_end_time - tn > 1.0e-10:
        if max_steps is not None and stats.steps >= max_steps:
            break
        tnp1 = tn + dt
        tcomp = time.perf_counter() - start
        k = eng.solve_step_dynamics(tnp1, dt)
        if k > params.max_iterations:
            dt = dt / 2
            stats.failed_steps += 1
            if dt < params.min_timestep:
                raise SimulationAborted(f"Δt < {params.min_timestep:.2e}: aborting simulation")
            eng.update_not_converged()
            successive = 0
            continue
        successive += 1
        stats.steps += 1
        stats.newton_iterations += k - 1
        stats.iterations_per_step.append(k - 1)
        if successive >= params.accelerate_after_success_it and k <= 8:
            dt = min(dt * 1.5, params.max_timestep)
            successive = 0
        if on_converged is not None:
            on_converged(step, tnp1, eng)
        if tcomp > params.tcompt_max and params.stop_long_simulation:
            break
        eng.update_converged()
        tn = tnp1
        step += 1
        if params.stop_on_energy_threshold:
            # the reference reads an undefined `energyⁿ⁺¹` here (run_simulation.jl:107); the
            # intended quantity, state.energyⁿ⁺¹.kinetic_energy, is used
            if eng.get("ENERGY")[1] < params.energy_threshold:
                break
    stats.t_end = tn
    stats.wall_s = time.perf_counter() - start
    return stats
