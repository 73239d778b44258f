"""Batches of independent structures (eb_create_batch): every instance of a batch follows, attempt
by attempt, the trajectory it has when run alone on the device with the batch's time steps, with
the same Newton iteration counts, and the oracle agrees. Instances differ in size (wires, beams
per wire) and geometry (seeded radius and pitch), each inside its own vessel SDF."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("U", "UDOT", "UDDOT", "W", "WDOT", "WDDOT", "R", "DR")


def stents():
    from endobeams_amd import configs
    shapes = [(4, 24, 0), (6, 32, 1), (8, 28, 2)]   # (wires, beams per wire, seed), one turn each
    return [configs.braided_stent(n_wires=w, el_per_wire=m, length=8e-3, sdf_n=48, seed=s) for (w, m, s) in shapes]


def batch_schedule(batch, attempts):
    """Runs the batch through the reference time loop; returns [(t, dt, accepted, k per instance)]."""
    from endobeams_amd import Stepper
    st = Stepper(batch, batch.params)
    st.initialize()
    out = []
    for _ in range(attempts):
        t, dt = st.tn + st.dt, st.dt
        ok = st.advance()
        out.append((t, dt, ok, batch.instance_iterations().copy()))
    return out


def replay(engine, schedule):
    engine.initialize_state_simulation()
    ks = []
    for (t, dt, ok, _) in schedule:
        ks.append(engine.solve_step_dynamics(t, dt))
        if ok:
            engine.update_converged()
        else:
            engine.update_not_converged()
    return ks


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_instances_follow_their_standalone_runs(precision):
    from endobeams_amd import HipBatchEngine, HipEngine, abi
    prec = abi.EB_PRECISION_FP64 if precision == "fp64" else abi.EB_PRECISION_FP32
    batch = HipBatchEngine(stents(), precision=prec)
    print(f"batch of {batch.S} instances, node offsets {batch.node_off.tolist()}", flush=True)
    sched = batch_schedule(batch, 6)
    print("attempts (dt, accepted, k per instance):", [(dt, ok, k.tolist()) for (_, dt, ok, k) in sched], flush=True)
    assert sum(ok for (_, _, ok, _) in sched) >= 3
    for s, (conf, params, inter) in enumerate(stents()):
        alone = HipEngine(conf, params, inter, precision=prec)
        ka = replay(alone, sched)
        assert ka == [int(k[s]) for (_, _, _, k) in sched], (s, ka)
        for f in FIELDS:
            a = alone.get(f)
            b = batch.get(f)[batch.instance_slice(f, s)]
            assert rel_err(b, a) <= 1e-13, (s, f, rel_err(b, a))
        alone.close()
    batch.close()


def test_large_fp32_batch_two_launch_element_kernel_follows_standalone_runs():
    """A precision-form batch of at least 2 048 element tiles (131 072 beams) runs its element
    kernel as two launches of two parts each (static and mass, then the Coriolis parts); a
    structure alone runs the four-part kernel. Instances of the batch follow their standalone runs
    attempt by attempt with the same Newton counts, as in the test above."""
    from endobeams_amd import HipBatchEngine, HipEngine, abi, configs

    def build(s):   # the braided stent of the benchmark (4 800 beams), at the time step it converges at
        return configs.braided_stent(sdf_n=64, seed=s, dt=1.25e-4)

    prec = abi.EB_PRECISION_FP32
    insts = [build(s) for s in range(28)]
    assert sum(len(c.beams) for (c, _, _) in insts) >= 64 * 2048
    batch = HipBatchEngine(insts, precision=prec)
    sched = batch_schedule(batch, 3)
    print("attempts (dt, accepted):", [(dt, ok) for (_, dt, ok, _) in sched], flush=True)
    assert sum(ok for (_, _, ok, _) in sched) >= 2
    for s in (0, 27):
        conf, params, inter = build(s)
        alone = HipEngine(conf, params, inter, precision=prec)
        ka = replay(alone, sched)
        assert ka == [int(k[s]) for (_, _, _, k) in sched], (s, ka)
        for f in FIELDS:
            a = alone.get(f)
            b = batch.get(f)[batch.instance_slice(f, s)]
            assert rel_err(b, a) <= 1e-13, (s, f, rel_err(b, a))
        alone.close()
    batch.close()


def test_batch_instance_matches_oracle():
    """One instance of a batch against the CPU restatement run alone (FP64)."""
    import oracle as O
    from endobeams_amd import HipBatchEngine
    batch = HipBatchEngine(stents())
    sched = batch_schedule(batch, 4)
    conf, params, inter = stents()[1]
    ref = O.OracleEngine(conf, params, inter)
    ka = replay(ref, sched)
    assert ka == [int(k[1]) for (_, _, _, k) in sched]
    # the converged kinematics to the tolerances of the single-structure trajectory parity test
    tol = {"U_N": 1e-10, "R_N": 1e-10, "UDOT_N": 1e-9}
    for f, t in tol.items():
        b = batch.get(f)[batch.instance_slice(f, 1)]
        e = rel_err(b, ref.get(f), floor=1e-30)
        print(f, e)
        assert e <= t, (f, e)
    batch.close()


def rings():
    from endobeams_amd import configs
    # rings of 24, 60 and 40 beams, starting near the plane: each is one closed chain whose
    # preconditioner carries the closure correction
    return [configs.ring_drop(n_el=n, height=h) for (n, h) in [(24, 0.06), (60, 0.5), (40, 0.06)]]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_of_rings_follows_the_standalone_rings(precision):
    from endobeams_amd import HipBatchEngine, HipEngine, abi
    prec = abi.EB_PRECISION_FP64 if precision == "fp64" else abi.EB_PRECISION_FP32
    batch = HipBatchEngine(rings(), precision=prec)
    sched = batch_schedule(batch, 6)
    print("attempts (dt, accepted, k per instance):", [(dt, ok, k.tolist()) for (_, dt, ok, k) in sched], flush=True)
    assert sum(ok for (_, _, ok, _) in sched) >= 3
    for s, (conf, params, inter) in enumerate(rings()):
        alone = HipEngine(conf, params, inter, precision=prec)
        ka = replay(alone, sched)
        assert ka == [int(k[s]) for (_, _, _, k) in sched], (s, ka)
        for f in FIELDS:
            a = alone.get(f)
            b = batch.get(f)[batch.instance_slice(f, s)]
            assert rel_err(b, a) <= 1e-13, (s, f, rel_err(b, a))
        alone.close()
    batch.close()


def nets():
    from endobeams_amd import configs
    # nets of 8×8, 12×12 and 10×10 crossings, started in contact with the sphere: each has a first
    # chain cover and a crossing cover, both per instance
    return [configs.net_on_sphere(n=n, sdf_n=48, gap=0.005) for n in (8, 12, 10)]


def test_batch_of_nets_follows_the_standalone_nets():
    from endobeams_amd import HipBatchEngine, HipEngine, abi
    batch = HipBatchEngine(nets(), precision=abi.EB_PRECISION_FP64)
    sched = batch_schedule(batch, 5)
    print("attempts (dt, accepted, k per instance):", [(dt, ok, k.tolist()) for (_, dt, ok, k) in sched], flush=True)
    assert sum(ok for (_, _, ok, _) in sched) >= 3
    for s, (conf, params, inter) in enumerate(nets()):
        alone = HipEngine(conf, params, inter)
        ka = replay(alone, sched)
        assert ka == [int(k[s]) for (_, _, _, k) in sched], (s, ka)
        for f in FIELDS:
            a = alone.get(f)
            b = batch.get(f)[batch.instance_slice(f, s)]
            assert rel_err(b, a) <= 1e-12, (s, f, rel_err(b, a))
        alone.close()
    batch.close()
