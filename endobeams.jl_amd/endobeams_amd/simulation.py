"""run_simulation — the host time loop of src/run_simulation.jl:1-119, driving an Engine.

The loop (adaptive Δt, rollback on Newton failure, abort below min_timestep, wall-clock stop) is
the reference's caller of the hot path; it stays on the host and calls the engine once per step.
`Stepper` exposes the same loop one converged step at a time (used by the benchmark).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

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


class Stepper:
    """State of run_simulation!'s time loop (run_simulation.jl:21-112) between steps."""

    def __init__(self, engine: Engine, params: SimulationParams):
        self.eng = engine
        self.p = params
        self.tn = 0.0
        self.dt = params.initial_timestep
        self.successive = 0
        self.step_index = 1
        self.stats = RunStats()
        self.start = time.perf_counter()
        self.stop = False

    def initialize(self):
        self.eng.initialize_state_simulation()

    @property
    def finished(self) -> bool:
        return not (self.p.simulation_end_time - self.tn > 1.0e-10)

    def advance(self, on_converged: Optional[Callable] = None) -> bool:
        """Attempts one step; returns True when it converged (and was committed)."""
        p = self.p
        tnp1 = self.tn + self.dt
        tcomp = time.perf_counter() - self.start
        k = self.eng.solve_step_dynamics(tnp1, self.dt)
        if k > p.max_iterations:
            self.dt = self.dt / 2
            self.stats.failed_steps += 1
            if self.dt < p.min_timestep:
                raise SimulationAborted(f"Δt < {p.min_timestep:.2e}: aborting simulation")
            self.eng.update_not_converged()
            self.successive = 0
            return False
        self.successive += 1
        self.stats.steps += 1
        self.stats.newton_iterations += k - 1
        self.stats.iterations_per_step.append(k - 1)
        if self.successive >= p.accelerate_after_success_it and k <= 8:
            self.dt = min(self.dt * 1.5, p.max_timestep)
            self.successive = 0
        if on_converged is not None:
            on_converged(self.step_index, tnp1, self.eng)
        self.stop = tcomp > p.tcompt_max and p.stop_long_simulation
        if self.stop:
            return True
        self.eng.update_converged()
        self.tn = tnp1
        self.step_index += 1
        if p.stop_on_energy_threshold:
            # the reference reads an undefined `energyⁿ⁺¹` here (run_simulation.jl:107); the
            # intended quantity, state.energyⁿ⁺¹.kinetic_energy, is used
            if self.eng.get("ENERGY")[1] < p.energy_threshold:
                self.stop = True
        return True


def run_simulation(conf: BeamsConfiguration, params: SimulationParams, inter: Optional[RigidInteraction] = None,
                   engine: Optional[Engine] = None, on_converged: Optional[Callable] = None,
                   max_steps: Optional[int] = None, driver: str = "host") -> RunStats:
    """run_simulation!(conf, params, inter) (src/run_simulation.jl:1-119).

    `engine` defaults to the HIP engine. `on_converged(step, t, engine)` is called after each
    converged step, before update_converged! (where the reference writes VTK output).
    `max_steps` bounds the number of converged steps (None: run to simulation_end_time).
    driver="device": the time loop runs inside the library (eb_run_steps), whole steps without a
    return to Python; `on_converged` is then not available."""
    eng = engine if engine is not None else HipEngine(conf, params, inter)
    if driver == "device":
        if on_converged is not None:
            raise ValueError("on_converged needs the host driver")
        start = time.perf_counter()
        eng.initialize_state_simulation()
        state = eng.run_steps(max_steps if max_steps is not None else 2**62)
        stats = RunStats(steps=state.steps, failed_steps=state.failed_steps,
                         newton_iterations=state.newton_iterations, t_end=state.t)
        stats.wall_s = time.perf_counter() - start
        return stats
    st = Stepper(eng, params)
    hook = on_converged
    if params.output_dir:
        # write_VTK of the initial state, then at the output cadence, before update_converged!
        # (run_simulation.jl:30-32, :86-90); the file carries tⁿ, as the reference passes it
        from .vtk import VTKDataBeams
        vtk = VTKDataBeams(conf, params.output_dir)
        vtk.write(0, st.step_index, 0.0, eng)
        out = {"counter": 1, "t": 0.0}

        def hook(step, tnp1, e):
            due = out["t"] + params.output_timestep
            if tnp1 > due or abs(tnp1 - due) <= 1.4901161193847656e-08 * max(abs(tnp1), abs(due)):
                vtk.write(out["counter"], step, st.tn, e)
                out["counter"] += 1
                out["t"] = tnp1
            if on_converged is not None:
                on_converged(step, tnp1, e)
    st.initialize()
    while not st.finished and not st.stop:
        if max_steps is not None and st.stats.steps >= max_steps:
            break
        st.advance(hook)
    st.stats.t_end = st.tn
    st.stats.wall_s = time.perf_counter() - st.start
    return st.stats
