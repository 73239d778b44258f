"""The quadratic B-spline prefilter of the signed distance field runs on the device at eb_create
(src/interactions/def_interactions.jl:67-82). Its coefficients, read back with
eb_get_sdf_coefficients, agree with the oracle's (which the CPU suite pins against scipy's
spline_filter) to the rounding of the two orders of operation in FP64 (the oracle divides by the
pivot, the device multiplies by its reciprocal), and to FP32 rounding in the precision form.
Coefficients passed prefiltered are kept as they are; a batch's instances hold the coefficients
they have alone."""
import numpy as np
import pytest

import oracle as O
from endobeams_amd import HipBatchEngine, HipEngine, abi, configs
from endobeams_amd.model import (BeamElementSurface, Beams, BeamsConfiguration, DiscreteSignedDistanceField,
                                 InteractionProperties, NodesBeams, RigidInteraction, SimulationParams)

pytestmark = pytest.mark.gpu


def setup(sdf):
    nodes = NodesBeams(np.array([[0.0, 0, 0], [1.0, 0, 0]]))
    beams = Beams(nodes, [[1, 2]], 1e5, 0.3, 1.0, 0.05, 0.0)
    conf = BeamsConfiguration(nodes, beams, None, None)
    inter = RigidInteraction(sdf, BeamElementSurface([1]), InteractionProperties(1.0, 0.1, 1.0, 0.0))
    return conf, SimulationParams.hht(-0.05), inter


def random_field(n, seed=0):
    rng = np.random.default_rng(seed)
    vals = rng.standard_normal((n[2], n[1], n[0]))
    return DiscreteSignedDistanceField(vals.reshape(-1), n, (0.3, 0.2, 0.25), (-1.0, 0.5, 2.0))


FIELDS = {
    "random_12x10x9": lambda: setup(random_field((12, 10, 9))),
    "random_1x7x33": lambda: setup(random_field((1, 7, 33), seed=1)),   # a single-sample axis
    "random_2x64x3": lambda: setup(random_field((2, 64, 3), seed=2)),
    "ring_plane_64x64x32": configs.ring_drop,
    "net_sphere_128": configs.net_on_sphere,
}


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", list(FIELDS))
def test_device_prefilter_matches_the_oracle(name):
    conf, params, inter = FIELDS[name]()
    ref = O.OracleEngine(conf, params, inter)
    n = inter.master.n
    want = ref.sdf_coefs(n)
    g64 = HipEngine(conf, params, inter)
    got = g64.sdf_coefficients()
    e64 = rel_max(got, want)
    g32 = HipEngine(conf, params, inter, precision=abi.EB_PRECISION_FP32)
    e32 = rel_max(g32.sdf_coefficients(), want)
    print(f"{name}: FP64 {e64:.2e}, FP32 {e32:.2e}", flush=True)
    assert e64 <= 1e-14
    assert e32 <= 1.2e-7
    g64.close()
    g32.close()


def test_prefiltered_coefficients_are_kept():
    conf, params, inter = FIELDS["random_12x10x9"]()
    ref = O.OracleEngine(conf, params, inter)
    sdf = inter.master
    coef = ref.sdf_coefs(sdf.n)
    pre = DiscreteSignedDistanceField(np.zeros(int(np.prod(sdf.n))), sdf.n, sdf.spacing, sdf.origin)
    pre.samples = coef.reshape(-1).copy()
    pre.prefiltered = True
    conf2, params2, inter2 = setup(pre)
    g = HipEngine(conf2, params2, inter2)
    assert np.array_equal(g.sdf_coefficients(), coef)
    g.close()


def test_batch_instances_hold_their_own_coefficients():
    inst = configs.stent_sweep(3, sdf_n=24)
    batch = HipBatchEngine(inst, precision=abi.EB_PRECISION_FP32)
    for s, (conf, params, inter) in enumerate(inst):
        alone = HipEngine(conf, params, inter, precision=abi.EB_PRECISION_FP32)
        assert np.array_equal(batch.sdf_coefficients(s), alone.sdf_coefficients()), s
        alone.close()
    batch.close()
