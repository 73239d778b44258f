"""Batches whose instances keep their own time step (eb_run_steps_batch) and whose linear-solve
failures stay with the failing instance.

run_simulation! (src/run_simulation.jl:38-112) halves Δt and repeats the step of the one
structure it runs when its Newton iteration needs more than max_iterations; it grows Δt ×1.5
after four easy steps. A batch of independent structures does that per instance: each instance
must follow, attempt by attempt, the trajectory it has alone. A linear solve that fails for one
instance (here: BiCGStab capped at one iteration, which the reference's net needs several of)
counts as that instance's Newton failure (k = max_iterations + 1) and must not stop the others.
Every kernel exits normally: the cap is a host-side iteration count.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import REPO, rel_err
from endobeams_amd import EBError, HipBatchEngine, HipEngine, abi, configs
from endobeams_amd.model import InteractionProperties, RigidInteraction

pytestmark = pytest.mark.gpu

FIELDS = ("U_N", "UDOT_N", "R_N", "TINT_N")


def cantilever(load, precision_tol=1e-8):
    conf, params, inter = configs.l_cantilever(tol=precision_tol, load=load)
    params.max_iterations = 8
    params.min_timestep = 1e-4
    return conf, params, inter


def states_of(engine, n):
    st = (abi.RunState * n)()
    for r in st:
        r.t, r.dt = 0.0, engine.params.initial_timestep
    return st


def alone(case, precision, steps):
    conf, params, inter = case
    e = HipEngine(conf, params, inter, precision=precision)
    e.initialize_state_simulation()
    state = e.run_steps(steps)
    return e, state


@pytest.mark.parametrize("precision", [abi.EB_PRECISION_FP64, abi.EB_PRECISION_FP32])
def test_batch_instances_keep_their_own_time_step(precision):
    """Three L-cantilevers with tip loads 5, 50 and 500 (max_iterations 8): the heavier loads
    halve Δt at different steps. In one batch each instance ends where it ends alone after as many
    converged steps: the same Newton iterations, time, Δt and repeated steps, and state, bit for
    bit."""
    loads = (5.0, 50.0, 500.0)
    steps = 12
    cases = [cantilever(f) for f in loads]
    b = HipBatchEngine(cases, precision=precision)
    b.initialize_state_simulation()
    st = b.run_steps_batch(steps, states_of(b, len(cases)))
    failed = []
    for s, case in enumerate(cases):
        r = st[s]
        # instances keep stepping until the slowest has its quota: some have more steps
        assert r.steps >= steps
        e, ref = alone(case, precision, r.steps)
        assert (r.steps, r.newton_iterations) == (ref.steps, ref.newton_iterations), s
        if r.last_k <= case[1].max_iterations:
            # the last attempt converged (no repeated attempt after the last converged step)
            assert (r.failed_steps, r.successive) == (ref.failed_steps, ref.successive), s
            assert r.t == ref.t and r.dt == ref.dt, (s, r.t, ref.t, r.dt, ref.dt)
        for f in FIELDS:
            np.testing.assert_array_equal(b.get(f)[b.instance_slice(f, s)], e.get(f), err_msg=f"{s} {f}")
        failed.append(r.failed_steps)
    # the test is only meaningful when the instances did not share one Δt history
    assert len(set(failed)) == len(loads), failed
    k = b.instance_iterations()
    assert k.shape == (3,)


def test_batch_instance_with_halvings_matches_the_oracle():
    """The hardest instance of the batch (tip load 500, several Δt halvings) against the CPU
    oracle's own time loop: identical attempt counts, state to 1e-10 (FP64)."""
    cases = [cantilever(f) for f in (5.0, 500.0)]
    steps = 8
    b = HipBatchEngine(cases, precision=abi.EB_PRECISION_FP64)
    b.initialize_state_simulation()
    st = b.run_steps_batch(steps, states_of(b, 2))
    conf, params, inter = cases[1]
    o = O.OracleEngine(conf, params, inter)
    o.initialize_state_simulation()
    ref = o.run_steps(steps)
    r = st[1]
    assert ref.failed_steps >= 2
    # the hard instance is the slowest: it ends at its quota, like the oracle's run
    assert (r.steps, r.failed_steps, r.newton_iterations) == (ref.steps, ref.failed_steps, ref.newton_iterations)
    assert r.t == pytest.approx(ref.t, rel=1e-15)
    for f in FIELDS:
        assert rel_err(b.get(f)[b.instance_slice(f, 1)], o.get(f), floor=1e-300) <= 1e-10, f


def test_batch_graph_replays_equal_eager_with_instance_time_steps():
    """The captured Newton iteration reads each instance's Δt from device memory: graph replays
    give the eager batch's trajectories bit for bit while the instances' Δt differ and change."""
    cases = [cantilever(f) for f in (5.0, 200.0, 500.0)]
    out = []
    for graph in (False, True):
        b = HipBatchEngine(cases, precision=abi.EB_PRECISION_FP64)
        b.set_exec_mode(graph)
        b.initialize_state_simulation()
        st = b.run_steps_batch(10, states_of(b, 3))
        out.append(([(r.t, r.dt, r.steps, r.failed_steps, r.newton_iterations) for r in st],
                    {f: b.get(f) for f in FIELDS}, b.exec_stats()))
    assert out[0][0] == out[1][0]
    for f in FIELDS:
        np.testing.assert_array_equal(out[0][1][f], out[1][1][f])
    assert out[1][2][2] > 0   # the graph was replayed


def wire_and_net():
    """The reference's sphere-contact wire (imposed displacement on dof 1) and the reference's
    net on its sphere, given the same contact properties so that they fit one batch."""
    props = InteractionProperties(1e-6, 0.01, 0.5, 1e-3, 1e-6, 1e-3, 0.0)
    wconf, wparams, winter = configs.reference_test_contact_2(1.0)
    winter = RigidInteraction(winter.master, winter.slave, props)
    pos = np.loadtxt(os.path.join(REPO, "tests", "golden", "pos_net.txt"))
    conn = np.loadtxt(os.path.join(REPO, "tests", "golden", "conn_net.txt"), dtype=np.int64)
    pos[:, 1] -= 0.0049
    nconf, nparams, ninter = configs.reference_net(pos, conn)
    for p in (wparams, nparams):
        p.initial_timestep, p.max_timestep, p.min_timestep = 1e-2, 1e-2, 1e-3
        p.tolerance_residual = p.tolerance_displacement = 1e-5
        p.max_iterations = 5
        p.krylov_maxit = 1
    return (wconf, wparams, winter), (nconf, nparams, ninter)


def test_a_failed_linear_solve_is_a_newton_failure_of_its_instance():
    """BiCGStab capped at one iteration: the open wire solves within it (its chain
    preconditioner is exact), the net does not. The net's attempts all report
    k = max_iterations + 1, its Δt halves until it falls below min_timestep and the net stops
    (stopped = 2); the wire advances as it does alone, bit for bit (the net's crossing chain
    cover holds no node of the wire, whose chain preconditioner is complete without it)."""
    wire, net = wire_and_net()
    b = HipBatchEngine([wire, net], precision=abi.EB_PRECISION_FP64)
    b.initialize_state_simulation()
    st = b.run_steps_batch(6, states_of(b, 2))
    w, n = st[0], st[1]
    assert n.stopped == 2 and n.steps == 0 and n.failed_steps == 4 and n.last_k == 6
    assert w.stopped == 0 and w.steps == 6 and w.failed_steps == 0
    e, ref = alone(wire, abi.EB_PRECISION_FP64, w.steps)
    assert (w.steps, w.newton_iterations) == (ref.steps, ref.newton_iterations) and w.t == ref.t
    for f in FIELDS:
        np.testing.assert_array_equal(b.get(f)[b.instance_slice(f, 0)], e.get(f), err_msg=f)
    # the net, alone, under the single-structure loop: the same failures end in the reference's abort
    conf, params, inter = net
    e = HipEngine(conf, params, inter)
    e.initialize_state_simulation()
    with pytest.raises(EBError, match="aborted"):
        e.run_steps(3)
