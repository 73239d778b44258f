"""eb_run_steps on the device (the time loop of src/run_simulation.jl:38-112 inside the library)
against the host loop (endobeams_amd.Stepper) driving the same library call by call: the same
attempts, Newton iterations and final state, bit for bit; for one structure and for a batch; and
the reference's abort when Δt falls below min_timestep."""
import numpy as np
import pytest

from endobeams_amd import EBError, HipBatchEngine, HipEngine, Stepper, abi, configs

pytestmark = pytest.mark.gpu

FIELDS = ("U_N", "UDOT_N", "R_N", "TINT_N", "ENERGY")


def host_run(eng, params, steps):
    st = Stepper(eng, params)
    st.initialize()
    while st.stats.steps < steps and not st.finished:
        st.advance()
    return st


def compare(make, params, steps):
    a = make()
    st = host_run(a, params, steps)
    b = make()
    b.initialize_state_simulation()
    state = b.run_steps(steps // 2)
    state = b.run_steps(steps - steps // 2, state)
    print(f"steps {state.steps}, halvings {state.failed_steps}, Newton {state.newton_iterations}, t {state.t:.3e}",
          flush=True)
    assert (state.steps, state.failed_steps, state.newton_iterations) == \
           (st.stats.steps, st.stats.failed_steps, st.stats.newton_iterations)
    assert state.t == st.tn and state.dt == st.dt
    for f in FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["l_cantilever", "contact_sphere", "stent"])
def test_device_time_loop_equals_host_loop(name):
    conf, params, inter = {"l_cantilever": configs.l_cantilever,
                           "contact_sphere": lambda: configs.reference_test_contact_2(0.1),
                           "stent": lambda: configs.braided_stent(n_wires=4, el_per_wire=40, sdf_n=64)}[name]()
    compare(lambda: HipEngine(conf, params, inter), params, 10)


def test_device_time_loop_batch():
    inst = configs.stent_sweep(3, sdf_n=32)
    compare(lambda: HipBatchEngine(inst, precision=abi.EB_PRECISION_FP32), inst[0][1], 6)


def test_abort_below_min_timestep():
    """The braided stent's first step at Δt = 1e-3 needs Δt halvings; with min_timestep = 6e-4
    the first halving aborts, as run_simulation! does."""
    conf, params, inter = configs.braided_stent(sdf_n=64)
    params.min_timestep = 6e-4
    eng = HipEngine(conf, params, inter)
    eng.initialize_state_simulation()
    with pytest.raises(EBError, match="aborted"):
        eng.run_steps(5)
    eng.close()
