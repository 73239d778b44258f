"""The OpenMP timing build of the oracle (liboracle_omp.so, the CPU baseline's all-core leg)
computes what the parity build computes: element and contact outputs in parallel, added in beam
order, with the same unfused arithmetic. A stent in its vessel, four steps, bit for bit."""
import os

import numpy as np

import oracle as O
from endobeams_amd import Stepper, configs


def test_timing_build_equals_parity_build():
    os.environ["OMP_NUM_THREADS"] = "4"
    out = []
    for timing in (False, True):
        conf, params, inter = configs.braided_stent(n_wires=4, el_per_wire=40, sdf_n=48)
        e = O.OracleEngine(conf, params, inter, timing=timing)
        st = Stepper(e, params)
        st.initialize()
        for _ in range(4):
            st.advance()
        out.append((st.stats.iterations_per_step, {f: e.get(f) for f in ("U_N", "R_N", "TC_N", "K_N")}))
    assert out[0][0] == out[1][0]
    for f in out[0][1]:
        np.testing.assert_array_equal(out[0][1][f], out[1][1][f], err_msg=f)
