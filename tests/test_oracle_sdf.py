"""Pins of the quadratic B-spline SDF (Interpolations.jl BSpline(Quadratic(Reflect(OnCell()))),
src/interactions/def_interactions.jl:67-82) against scipy.ndimage, an independent implementation
of the same spline: spline_filter(order=2, mode='reflect') solves the same (⅛, ¾, ⅛) system with
half-sample symmetric ends (c₀ = c₁, c_{n+1} = c_n), and map_coordinates evaluates it."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from endobeams_amd.model import (BeamElementSurface, Beams, BeamsConfiguration, DiscreteSignedDistanceField,
                                 InteractionProperties, NodesBeams, RigidInteraction, SimulationParams)


def engine(sdf):
    nodes = NodesBeams(np.array([[0.0, 0, 0], [1.0, 0, 0]]))
    beams = Beams(nodes, [[1, 2]], 1e5, 0.3, 1.0, 0.05, 0.0)
    conf = BeamsConfiguration(nodes, beams, None, None)
    inter = RigidInteraction(sdf, BeamElementSurface([1]), InteractionProperties(1.0, 0.1, 1.0, 0.0))
    return O.OracleEngine(conf, SimulationParams.hht(-0.05), inter)


@pytest.fixture(scope="module")
def field():
    rng = np.random.default_rng(0)
    n = (12, 10, 9)  # scipy's recursive filter start is exact to 1e-15 only for lines of >= 9 samples
    vals = rng.standard_normal((n[2], n[1], n[0]))  # z, y, x (x fastest in memory)
    sdf = DiscreteSignedDistanceField(vals.reshape(-1), n, (0.3, 0.2, 0.25), (-1.0, 0.5, 2.0))
    return vals, sdf, engine(sdf)


def test_prefilter_matches_scipy(field):
    vals, sdf, eng = field
    coef = eng.sdf_coefs(sdf.n)                        # (nz+2, ny+2, nx+2) padded
    ref = ndimage.spline_filter(vals, order=2, mode="reflect")
    np.testing.assert_allclose(coef[1:-1, 1:-1, 1:-1], ref, atol=1e-12, rtol=0)
    # reflecting ends: c₀ = c₁, c_{n+1} = c_n along every axis
    np.testing.assert_array_equal(coef[:, :, 0], coef[:, :, 1])
    np.testing.assert_array_equal(coef[:, :, -1], coef[:, :, -2])
    np.testing.assert_array_equal(coef[0], coef[1])
    np.testing.assert_array_equal(coef[:, -1], coef[:, -2])


def test_values_match_scipy_and_interpolate(field):
    vals, sdf, eng = field
    rng = np.random.default_rng(1)
    d, x0 = np.array(sdf.spacing), np.array(sdf.origin)
    for _ in range(40):
        idx = rng.uniform(0, 1, 3) * (np.array(sdf.n) - 1)
        p = x0 + idx * d
        v, g, H = eng.sdf_eval(p)
        ref = ndimage.map_coordinates(vals, [[idx[2]], [idx[1]], [idx[0]]], order=2, mode="reflect")[0]
        assert v == pytest.approx(ref, abs=1e-12)
    # interpolation condition at the samples
    for (i, j, k) in [(0, 0, 0), (3, 2, 1), (11, 9, 8)]:
        p = x0 + np.array([i, j, k]) * d
        assert eng.sdf_eval(p)[0] == pytest.approx(vals[k, j, i], abs=1e-12)


def test_gradient_and_hessian_are_derivatives(field):
    vals, sdf, eng = field
    d, x0 = np.array(sdf.spacing), np.array(sdf.origin)
    p = x0 + np.array([3.3, 2.6, 2.2]) * d     # away from knot-switch planes
    v, g, H = eng.sdf_eval(p)
    h = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        vp, gp, _ = eng.sdf_eval(p + e)
        vm, gm, _ = eng.sdf_eval(p - e)
        assert (vp - vm) / (2 * h) == pytest.approx(g[a], rel=1e-7, abs=1e-7)
        np.testing.assert_allclose((gp - gm) / (2 * h), H[:, a], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(H, H.T, atol=1e-12)


def test_linear_field_is_reproduced_in_the_interior():
    n = (30, 30, 30)
    ax = [np.arange(m) * 0.1 for m in n]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    vals = 0.3 * X - 0.7 * Y + 1.1 * Z + 0.2
    eng = engine(DiscreteSignedDistanceField(vals.reshape(-1), n, (0.1,) * 3, (0.0,) * 3))
    p = np.array([1.43, 1.57, 1.61])
    v, g, H = eng.sdf_eval(p)
    assert v == pytest.approx(0.3 * p[0] - 0.7 * p[1] + 1.1 * p[2] + 0.2, abs=1e-10)
    np.testing.assert_allclose(g, [0.3, -0.7, 1.1], atol=1e-9)
    np.testing.assert_allclose(H, np.zeros((3, 3)), atol=1e-7)


def test_prefiltered_coefficients_are_accepted(field):
    """eb_model_desc.sdf_prefiltered: the padded coefficients themselves (as the reference's
    interpolant stores them) give the same surface as the raw samples."""
    vals, sdf, eng = field
    coef = eng.sdf_coefs(sdf.n)
    pre = DiscreteSignedDistanceField(np.zeros(int(np.prod(sdf.n))), sdf.n, sdf.spacing, sdf.origin)
    pre.samples = coef.reshape(-1).copy()
    pre.prefiltered = True
    eng2 = engine(pre)
    p = np.array(sdf.origin) + np.array([2.5, 3.1, 1.7]) * np.array(sdf.spacing)
    a, b = eng.sdf_eval(p), eng2.sdf_eval(p)
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[1], b[1])
