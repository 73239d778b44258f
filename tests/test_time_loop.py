"""eb_run_steps, run_simulation!'s time loop inside the library (src/run_simulation.jl:38-112),
against the host restatement of the same loop (endobeams_amd.Stepper) on the oracle: the same
attempts (Δt halvings and ×1.5 growth), the same Newton iterations, the same final state, bit for
bit. Loads and imposed displacements come through the step callback."""
import numpy as np
import pytest

import oracle as O
from endobeams_amd import Stepper, configs

FIELDS = ("U_N", "UDOT_N", "R_N", "TINT_N", "ENERGY")

CASES = {
    # tip load ramp 50 t on the free end; Δt = 0.25 fixed
    "l_cantilever": lambda: configs.l_cantilever(),
    # imposed triangle-wave displacement on dof 1, wire against a sphere (test/test_beams2rigid_contact_2_analytical.jl)
    "contact_sphere": lambda: configs.reference_test_contact_2(0.1),
    # load ramp at the elbow, Δt fixed (test/test_beams.jl)
    "test_beams": lambda: configs.reference_test_beams(0.01),
    # the same, started at Δt = 2e-5 below max_timestep = 1e-4: the ×1.5 growth after easy steps
    "test_beams_growing_dt": lambda: growing(configs.reference_test_beams(0.01)),
}


def growing(case):
    conf, params, inter = case
    params.initial_timestep, params.max_timestep, params.min_timestep = 2e-5, 1e-4, 1e-6
    return conf, params, inter


@pytest.mark.parametrize("name", list(CASES))
def test_library_time_loop_equals_host_loop(name):
    conf, params, inter = CASES[name]()
    steps = 12
    host = O.OracleEngine(conf, params, inter)
    st = Stepper(host, params)
    st.initialize()
    while st.stats.steps < steps and not st.finished:
        st.advance()
    dev = O.OracleEngine(conf, params, inter)
    dev.initialize_state_simulation()
    # two chunks, to exercise continuing from a returned state
    state = dev.run_steps(5)
    state = dev.run_steps(steps - 5, state)
    assert (state.steps, state.failed_steps, state.newton_iterations) == \
           (st.stats.steps, st.stats.failed_steps, st.stats.newton_iterations)
    assert state.t == st.tn and state.dt == st.dt and state.successive == st.successive
    if name == "test_beams_growing_dt":
        assert state.dt > params.initial_timestep
    for f in FIELDS:
        assert np.array_equal(dev.get(f), host.get(f)), f


def test_time_loop_stops_at_the_end_time():
    conf, params, inter = configs.reference_test_beams(3e-4)   # 3 steps of Δt = 1e-4
    eng = O.OracleEngine(conf, params, inter)
    eng.initialize_state_simulation()
    state = eng.run_steps(100)
    assert state.steps == 3 and abs(state.t - 3e-4) < 1e-12
    again = eng.run_steps(100, state)
    assert again.steps == 3


def test_callback_errors_propagate():
    conf, params, inter = CASES["l_cantilever"]()
    eng = O.OracleEngine(conf, params, inter)
    eng.initialize_state_simulation()

    def broken(t):
        raise RuntimeError("load table ends")
    eng.update_loads = broken
    with pytest.raises(RuntimeError, match="load table ends"):
        eng.run_steps(3)
