"""The one-workgroup Newton iteration of a small structure (eb_set_exec_mode(EB_EXEC_FUSED)): the
predictor (src/predictor.jl:1-27) as one launch and each iteration of the corrector
(src/corrector.jl:15-53) as the element kernel plus one single-workgroup launch running the same
device bodies the eagerly launched path uses, iterations launched in batches without host round
trips. Every trajectory, Newton iteration count, energy and matrix is the eager path's, bit for bit
(which the parity suite checks against the oracle); one case is also checked against the oracle
directly. Contexts the mode does not cover are refused."""
import numpy as np
import pytest

from conftest import rel_err
from endobeams_amd import EBError, HipBatchEngine, HipEngine, Stepper, abi, configs
import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("U", "UDOT", "UDDOT", "W", "WDOT", "WDDOT", "R", "DR", "TINT", "TK", "TC", "K_N", "C_N", "M_N", "ENERGY")


def run(engine, params, steps, fields=FIELDS):
    st = Stepper(engine, params)
    st.initialize()
    attempts = 0
    while st.stats.steps < steps and attempts < 4 * steps and not st.finished:
        attempts += 1
        st.advance()
    return list(st.stats.iterations_per_step), st.stats.failed_steps, {f: engine.get(f).copy() for f in fields}


CASES = {
    # L-cantilever (20 beams, one open chain, clamped node, ramped tip load): Δt changes as the load ramps
    "cantilever_fp64": lambda: (configs.l_cantilever(), abi.EB_PRECISION_FP64, 25),
    "cantilever_fp32": lambda: (configs.l_cantilever(), abi.EB_PRECISION_FP32, 25),
    # ring on an SDF plane (60 beams, a closed chain: the Woodbury closure; contact and friction)
    "ring_fp64": lambda: (configs.ring_drop(), abi.EB_PRECISION_FP64, 30),
    "ring_fp32": lambda: (configs.ring_drop(), abi.EB_PRECISION_FP32, 30),
    # the reference's wire on a sphere: imposed displacement (the lift) and analytical contact
    "wire_sphere_fp64": lambda: (configs.reference_test_contact_2(), abi.EB_PRECISION_FP64, 20),
}


@pytest.mark.parametrize("case", list(CASES))
def test_fused_step_equals_eager_launches(case):
    (conf, params, inter), prec, steps = CASES[case]()
    eager = HipEngine(conf, params, inter, precision=prec)
    k0, f0_fail, f0 = run(eager, params, steps)
    eager.close()
    (conf, params, inter), prec, steps = CASES[case]()
    fused = HipEngine(conf, params, inter, precision=prec)
    fused.set_exec_mode("fused")
    k1, f1_fail, f1 = run(fused, params, steps)
    assert fused.exec_stats()[0] == abi.EB_EXEC_FUSED
    assert k1 == k0 and f1_fail == f0_fail, (k0, k1)
    for f in FIELDS:
        assert np.array_equal(f1[f], f0[f]), (case, f, rel_err(f1[f], f0[f]))
    fused.close()


def test_fused_cantilever_against_the_oracle():
    """The fused FP64 path against the C++ FP64 restatement over 25 steps: the same Newton
    iteration counts, positions and rotations to 1e-10."""
    conf, params, inter = configs.l_cantilever()
    g = HipEngine(conf, params, inter)
    g.set_exec_mode("fused")
    kg, _, fg = run(g, params, 25)
    conf, params, inter = configs.l_cantilever()
    o = O.OracleEngine(conf, params, inter)
    ko, _, fo = run(o, params, 25, ("U", "R", "UDOT", "TINT"))
    assert kg == ko
    for f in ("U", "R", "UDOT", "TINT"):
        assert rel_err(fg[f], fo[f]) < 1e-10, f


def test_fused_refuses_what_it_does_not_cover():
    """A net (crossing couplings no chain holds) and a batch of two structures keep the eager path."""
    import os
    here = os.path.join(os.path.dirname(__file__), "golden")
    pos = np.loadtxt(os.path.join(here, "pos_net.txt"))
    conn = np.loadtxt(os.path.join(here, "conn_net.txt"), dtype=int)
    net = HipEngine(*configs.reference_net(pos, conn))
    with pytest.raises(EBError):
        net.set_exec_mode("fused")
    assert net.exec_stats()[0] == abi.EB_EXEC_EAGER
    b = HipBatchEngine([configs.l_cantilever(), configs.l_cantilever(load=5.0)])
    with pytest.raises(EBError):
        b.set_exec_mode("fused")
