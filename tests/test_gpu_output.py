"""eb_output_points (recompute_at_gausspts_beams!, src/IO/visualization_beams.jl:132-181) on the
device against the oracle on the same randomised state, FP64, 1e-10 norm-wise; and after a few
steps of a simulation."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from conftest import rel_err
from endobeams_amd import HipEngine, Stepper, abi, configs

pytestmark = pytest.mark.gpu


def randomize(engines, conf, seed):
    rng = np.random.default_rng(seed)
    N = len(conf.nodes)
    u = rng.standard_normal(3 * N) * 1e-3
    ud = rng.standard_normal(3 * N) * 1e-2
    R = np.transpose(Rotation.from_rotvec(rng.standard_normal((N, 3)) * 0.05).as_matrix(), (0, 2, 1)).reshape(-1)
    for e in engines:
        e.set("U", u)
        e.set("UDOT", ud)
        e.set("R", R)


@pytest.mark.parametrize("name", ["ring", "stent", "net"])
@pytest.mark.parametrize("n_points", [10, 3])
def test_output_points_match_the_oracle(name, n_points):
    conf, params, inter = {"ring": configs.ring_drop,
                           "stent": lambda: configs.braided_stent(n_wires=4, el_per_wire=40, sdf_n=32),
                           "net": lambda: configs.net_on_sphere(n=12, sdf_n=32)}[name]()
    g, o = HipEngine(conf, params, inter), O.OracleEngine(conf, params, inter)
    randomize([g, o], conf, 5)
    a, b = g.output_points(n_points), o.output_points(n_points)
    for k in ("position", "displacement", "velocity", "strain", "stress"):
        e = rel_err(a[k], b[k], floor=1e-30)
        assert e <= 1e-10, (k, e)
    g.close()


def test_output_points_after_steps_fp32_form():
    conf, params, inter = configs.ring_drop()
    g = HipEngine(conf, params, inter, precision=abi.EB_PRECISION_FP32)
    o = O.OracleEngine(conf, params, inter)
    for e in (g, o):
        st = Stepper(e, params)
        st.initialize()
        for _ in range(5):
            st.advance()
    a, b = g.output_points(), o.output_points()
    assert rel_err(a["position"], b["position"]) <= 1e-8
    g.close()
