"""The roofline's algorithmic flop figure for the element kernel is the operation count of the
oracle's restatement of compute_beams (counting build, oracle/flopcount.cpp): this keeps the
constant in bench.py equal to it. The count without the operations on exact zeros (structural
zeros of the dense 3x3 algebra, which the element kernel's build folds away) is reported beside it."""
import ctypes as C
import importlib.util
import os

import numpy as np

import oracle as O
from conftest import REPO
from test_oracle_element import GW, GZ, Elem


def count(el):
    lib = O.load()
    lib.ebo_count_beam.argtypes = [C.POINTER(C.c_double)] * 2 + [C.c_int] + [C.POINTER(C.c_double)] * 2 + \
        [C.POINTER(C.c_longlong)]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    e, p = el.pack(), np.ascontiguousarray(el.props)
    w, z = np.array(GW), np.array(GZ)
    out = np.zeros(4, np.int64)
    lib.ebo_count_beam(dp(e), dp(p), 3, dp(w), dp(z), out.ctypes.data_as(C.POINTER(C.c_longlong)))
    return out


def test_bench_flop_constant_matches_the_counting_build():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    counts = [count(Elem(s)) for s in range(3)]
    # the DAG has no data-dependent branches once rotations are non-trivial: same count every time
    assert all((c == counts[0]).all() for c in counts)
    assert counts[0][0] == bench.FLOPS_PER_BEAM
    assert counts[0][3] == bench.FLOPS_PER_BEAM_NONZERO
    assert counts[0][3] < counts[0][0]
    assert counts[0][1] == 8 and counts[0][2] == 16  # sqrt; acos/sin/cos
    # damping adds the damping forces and damping·M to Cᵏ
    assert count(Elem(0, damping=0.4))[0] > bench.FLOPS_PER_BEAM
