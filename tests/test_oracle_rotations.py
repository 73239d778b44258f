"""Closed-form pins of the oracle's rotation utilities (src/beams/utils_beams.jl, def_beam.jl:93-116)."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O

EPS = np.finfo(float).eps


def test_rotation_matrix_is_rodrigues():
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.standard_normal(3) * rng.uniform(1e-6, 3.0)
        R = O.rotation_matrix(v)
        np.testing.assert_allclose(R, Rotation.from_rotvec(v).as_matrix(), rtol=0, atol=1e-14)
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-14)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)
    np.testing.assert_array_equal(O.rotation_matrix(np.array([5 * EPS, 0, 0])), np.eye(3))


def test_toangle_inverts_rotation_matrix():
    rng = np.random.default_rng(1)
    for _ in range(50):
        axis = rng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        th = rng.uniform(1e-3, np.pi - 1e-3)
        v = th * axis
        np.testing.assert_allclose(O.toangle(O.rotation_matrix(v)), v, rtol=0, atol=1e-10 + 1e-12 / th)
    # identity → the 10ε floor times a zero axis
    np.testing.assert_array_equal(O.toangle(np.eye(3)), np.zeros(3))


def test_Ts_inverse_times_Ts_is_identity():
    rng = np.random.default_rng(2)
    for _ in range(30):
        v = rng.standard_normal(3) * rng.uniform(1e-4, 2.5)
        np.testing.assert_allclose(O.Ts_inv(v) @ O.Ts(v), np.eye(3), atol=1e-12)
    np.testing.assert_array_equal(O.Ts_inv(np.zeros(3)), np.eye(3))


def test_Ts_is_the_rotation_jacobian():
    """Tₛ(Θ) maps δΘ to the spatial spin: exp(Θ+δΘ) ≈ exp(Tₛ(Θ)δΘ) exp(Θ)."""
    rng = np.random.default_rng(3)
    v = rng.standard_normal(3)
    h = 1e-6
    T = O.Ts(v)
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        W = (O.rotation_matrix(v + d) - O.rotation_matrix(v - d)) / (2 * h) @ O.rotation_matrix(v).T
        w = np.array([W[2, 1], W[0, 2], W[1, 0]])
        np.testing.assert_allclose(w, T[:, k], atol=1e-8)


def test_eta_mu_series_limits():
    e, m = O.eta_mu(np.zeros(3))
    assert e == 1 / 12 and m == 1 / 360
    # continuity of the closed forms towards the series values (the reference's μ formula cancels
    # catastrophically below θ ≈ 1e-2 and is reproduced as written, so continuity is read at 0.05)
    e, m = O.eta_mu(np.array([1e-3, 0, 0]))
    assert e == pytest.approx(1 / 12, rel=1e-5)
    e, m = O.eta_mu(np.array([0.05, 0, 0]))
    assert m == pytest.approx(1 / 360, rel=1e-2)
    # η(θ) = (2 sin θ - θ(1 + cos θ)) / (2θ² sin θ) at θ = 1
    e, m = O.eta_mu(np.array([0.6, 0.8, 0.0]))
    assert e == pytest.approx((2 * np.sin(1) - (1 + np.cos(1))) / (2 * np.sin(1)), rel=1e-14)
    assert m == pytest.approx((1 + np.sin(1) - 8 * np.sin(0.5) ** 2) / (4 * np.sin(0.5) ** 2), rel=1e-13)


def test_get_Re0_maps_e1_to_chord():
    rng = np.random.default_rng(4)
    for _ in range(20):
        X1, X2 = rng.standard_normal(3), rng.standard_normal(3)
        R = O.get_Re0(X1, X2)
        np.testing.assert_allclose(R[:, 0], (X2 - X1) / np.linalg.norm(X2 - X1), atol=1e-14)
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-14)
    np.testing.assert_array_equal(O.get_Re0(np.zeros(3), np.array([2.0, 0, 0])), np.eye(3))
    np.testing.assert_array_equal(O.get_Re0(np.zeros(3), np.array([-2.0, 0, 0])), np.diag([-1.0, 1.0, -1.0]))
