"""Pins the oracle end to end on the only known answers the reference holds:
test/test_beams.jl:113-119 (tip uⁿ[3] of the L-frame at t = 0.15, 1.5 and 15, atol 1e-2),
with that script's exact setup (:12-104)."""
import numpy as np
import pytest

import oracle as O
from endobeams_amd import configs, run_simulation


@pytest.mark.parametrize("t_end,expected", [(0.15, -0.0100), (1.5, 5.8332), (15.0, 10.3849)])
def test_reference_known_answers(t_end, expected):
    conf, params, inter = configs.reference_test_beams(t_end)
    eng = O.OracleEngine(conf, params, inter)
    stats = run_simulation(conf, params, inter, engine=eng)
    tip = eng.get("U_N").reshape(-1, 3)[-1]
    assert stats.failed_steps == 0
    assert abs(stats.t_end - t_end) < 1e-9
    assert tip[2] == pytest.approx(expected, abs=1e-2)


def test_reference_contact_script_runs():
    """test/test_beams2rigid_contact_2_analytical.jl has no assertion; the oracle must at least
    run its setup (imposed triangle wave, frictionless sphere contact) without aborting, and the
    wire must come into contact (non-zero contact energy) without crossing the sphere."""
    conf, params, inter = configs.reference_test_contact_2(1.5)
    eng = O.OracleEngine(conf, params, inter)
    seen = {"gap": np.inf}
    c = np.array(inter.master.center)
    X0 = conf.nodes.X0

    def watch(step, t, e):
        x = X0 + e.get("U").reshape(-1, 3)
        seen["gap"] = min(seen["gap"], float(np.min(np.linalg.norm(x - c, axis=1)) - inter.master.radius))

    stats = run_simulation(conf, params, inter, engine=eng, on_converged=watch)
    # Δt halves on non-converged steps (max_iterations = 5), so the last step may overshoot t_end
    assert 1.5 - 1e-10 <= stats.t_end < 1.5 + params.max_timestep
    assert eng.get("ENERGY")[2] > 0.0
    # kₙ = 10 is a soft penalty: the driven wire indents the sphere by ~1.5 mm but is stopped well
    # before crossing it (the sphere radius is 40 mm)
    assert -0.005 < seen["gap"] < 0.0


def test_reference_net_fixture_runs_on_the_oracle():
    """The reference's net (tests/golden, from test/input/) in its contact-test setup, lowered onto
    the sphere: the oracle converges every step and the contact forces are active."""
    import os
    import numpy as np
    import oracle as O
    from conftest import REPO
    from endobeams_amd import configs, run_simulation
    pos = np.loadtxt(os.path.join(REPO, "tests", "golden", "pos_net.txt"))
    conn = np.loadtxt(os.path.join(REPO, "tests", "golden", "conn_net.txt"), dtype=np.int64)
    assert pos.shape == (667, 3) and conn.shape == (786, 2) and conn.min() == 1 and conn.max() == 667
    pos[:, 1] -= 0.0049
    conf, params, inter = configs.reference_net(pos, conn)
    o = O.OracleEngine(conf, params, inter)
    st = run_simulation(conf, params, inter, engine=o, max_steps=5)
    assert st.steps == 5 and st.failed_steps == 0
    assert np.linalg.norm(o.get("TC_N")) > 0
