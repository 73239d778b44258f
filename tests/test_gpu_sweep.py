"""The stent design sweep (512 braided stents of uneven sizes, each in its own vessel SDF on a
128³ FP32 grid, advanced together with the FP32 element matrices and the learned warm start) on
its extreme instances: the smallest draw (instance 301: 12 wires × 106 beams = 1 272 beams), the
largest (instance 65: 48 × 300 = 14 400) and the median (instance 203: 5 376).

* In one FP32 batch, with a network trained by the library's own kernels installed (evaluated at
  every prediction), each instance ends where it ends alone with the same network: the same
  Newton iterations and the same state, bit for bit.
* In an FP64 batch, the smallest instance follows the CPU oracle's own time loop (FP64): the same
  attempts and Newton iterations, positions and rotations within 1e-10.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err
from endobeams_amd import HipBatchEngine, HipEngine, abi, configs
from endobeams_amd import learned as L

pytestmark = pytest.mark.gpu

SMALLEST, LARGEST, MEDIAN = 301, 65, 203
FIELDS = ("U_N", "UDOT_N", "R_N", "TINT_N", "TC_N")


def instance(i):
    return configs.stent_sweep_instance(i, sdf_n=128)


def test_sweep_extremes_instance_sizes():
    sizes = {i: len(instance(i)[0].beams) for i in (SMALLEST, LARGEST, MEDIAN)}
    assert sizes == {SMALLEST: 1272, LARGEST: 14400, MEDIAN: 5376}


def trained_network():
    tr = [configs.braided_stent(n_wires=8, el_per_wire=60, sdf_n=64, seed=2000 + s) for s in range(2)]
    eng = HipBatchEngine(tr, precision=abi.EB_PRECISION_FP32)
    X, T = L.collect(eng, tr[0][1], 8)
    eng.close()
    return L.WarmStart.train(X, T, epochs=3, batch=4096, lr=1e-3, seed=0)


def test_sweep_batch_with_network_equals_each_instance_alone():
    ws = trained_network()
    cases = [instance(i) for i in (SMALLEST, LARGEST, MEDIAN)]
    b = HipBatchEngine(cases, precision=abi.EB_PRECISION_FP32)
    ws.install(b)
    L.set_backoff(b, 0)
    b.initialize_state_simulation()
    st = b.run_steps_batch(6)
    stats = L.predictor_stats(b)
    assert stats[0] + stats[1] > 0          # the network's start was evaluated against the reference's
    for s, case in enumerate(cases):
        r = st[s]
        e = HipEngine(*case, precision=abi.EB_PRECISION_FP32)
        ws.install(e)
        L.set_backoff(e, 0)
        e.initialize_state_simulation()
        ref = e.run_steps(r.steps)
        assert (r.steps, r.newton_iterations) == (ref.steps, ref.newton_iterations), s
        for f in FIELDS:
            np.testing.assert_array_equal(b.get(f)[b.instance_slice(f, s)], e.get(f), err_msg=f"{s} {f}")
        e.close()


def test_sweep_smallest_instance_in_a_batch_matches_the_oracle():
    cases = [instance(SMALLEST), instance(LARGEST)]
    b = HipBatchEngine(cases, precision=abi.EB_PRECISION_FP64)
    b.initialize_state_simulation()
    st = b.run_steps_batch(4)
    r = st[0]
    o = O.OracleEngine(*instance(SMALLEST))
    o.initialize_state_simulation()
    ref = o.run_steps(r.steps)
    assert (r.steps, r.newton_iterations) == (ref.steps, ref.newton_iterations)
    if r.last_k <= cases[0][1].max_iterations:
        assert r.failed_steps == ref.failed_steps and r.t == pytest.approx(ref.t, rel=1e-15)
    for f in ("U_N", "R_N"):
        assert rel_err(b.get(f)[b.instance_slice(f, 0)], o.get(f), floor=1e-300) <= 1e-10, f
    assert rel_err(b.get("UDOT_N")[b.instance_slice("UDOT_N", 0)], o.get("UDOT_N"), floor=1e-300) <= 1e-9
