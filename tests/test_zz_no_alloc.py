"""eb_output_points_noalloc, the entry point that allocates nothing (staging in a fixed device
array), against eb_output_points (staging allocated per call) through the C ABI: identical
points and fields on two calls in a row, the second on a smaller structure with fewer points
(the array keeps the first call's contents, which must not show); a request beyond the array's
size is refused with EB_ERR_ARG before any launch."""
import ctypes as C

import numpy as np
import pytest

from endobeams_amd import HipEngine, abi, configs

pytestmark = pytest.mark.gpu

FIELDS = ("position", "displacement", "velocity", "strain", "stress")


def noalloc(e, n_points, which=FIELDS):
    m = e.sizes.n_beams * n_points
    out = {f: (np.full((m, 3), np.nan) if f in ("position", "displacement", "velocity") else np.full(m, np.nan))
           for f in which}
    ptr = lambda f: out[f].ctypes.data_as(C.POINTER(C.c_double)) if f in out else None
    st = e.fn["output_points_noalloc"](e.h, int(n_points), ptr("position"), ptr("displacement"), ptr("velocity"),
                                       ptr("strain"), ptr("stress"))
    return st, out


def moved(case):
    """A structure after a few time steps (non-trivial displacements, velocities, rotations)."""
    conf, params, inter = case
    e = HipEngine(conf, params, inter)
    e.initialize_state_simulation()
    e.run_steps(3)
    return e


def test_noalloc_output_equals_allocating_output_on_two_calls():
    big = moved(configs.braided_stent(n_wires=8, el_per_wire=60, sdf_n=48))     # 480 beams
    small = moved(configs.ring_drop(n_el=24))                                     # 24 beams
    for e, n_points in ((big, 10), (small, 4)):
        want = e.output_points(n_points)
        st, got = noalloc(e, n_points)
        assert st == abi.EB_OK, e.fn["last_error"](e.h)
        for f in FIELDS:
            np.testing.assert_array_equal(got[f], want[f], err_msg=f)
    # a subset of the fields: the others are neither written nor needed
    st, got = noalloc(small, 3, which=("strain", "velocity"))
    want = small.output_points(3)
    assert st == abi.EB_OK
    np.testing.assert_array_equal(got["strain"], want["strain"])
    np.testing.assert_array_equal(got["velocity"], want["velocity"])


def test_noalloc_refuses_requests_beyond_its_array():
    e = moved(configs.ring_drop(n_el=24))
    n_points = abi.EB_NOALLOC_SCRATCH_DOUBLES // (11 * 24) + 1      # all five fields: 11 doubles a point
    st = e.fn["output_points_noalloc"](e.h, n_points, None, None, None, None, None)
    assert st == abi.EB_OK                                           # nothing requested: nothing staged
    pos = np.empty((24 * n_points, 3))
    big = [np.empty((24 * n_points, 3)) for _ in range(2)] + [np.empty(24 * n_points) for _ in range(2)]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    st = e.fn["output_points_noalloc"](e.h, n_points, dp(pos), dp(big[0]), dp(big[1]), dp(big[2]), dp(big[3]))
    assert st == abi.EB_ERR_ARG
    assert b"exceed the staging" in e.fn["last_error"](e.h)
    # the context is unharmed: a valid call still matches the allocating entry point
    st, got = noalloc(e, 5)
    assert st == abi.EB_OK
    np.testing.assert_array_equal(got["position"], e.output_points(5)["position"])
