"""Pins of the oracle's contact path against closed forms and finite differences:
gap functions (utils_compute_interactions_beams.jl), regularisation
(interactions_beams_regularization.jl) and compute_contact_beams, whose Kᶜ and Cᶜ must be the
derivatives of Tᶜ with respect to the configuration (spatial spins) and the nodal velocities."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from endobeams_amd import configs
from endobeams_amd.model import (BeamElementSurface, Beams, BeamsConfiguration, DiscreteSignedDistanceField,
                                 InteractionProperties, NodesBeams, RigidInteraction, SimulationParams, SphereSurface,
                                 beam_material_properties, get_Re0)

GW = (5 / 9, 8 / 9, 5 / 9)
GZ = (-np.sqrt(3 / 5), 0.0, np.sqrt(3 / 5))


def expm(v):
    return Rotation.from_rotvec(v).as_matrix()


def _engine(master, props):
    nodes = NodesBeams(np.array([[0.0, 0, 0], [1.0, 0, 0]]))
    beams = Beams(nodes, [[1, 2]], 1e5, 0.3, 1.0, 0.05, 0.0)
    conf = BeamsConfiguration(nodes, beams, None, None)
    inter = RigidInteraction(master, BeamElementSurface([1]), props)
    return O.OracleEngine(conf, SimulationParams.hht(-0.05), inter)


def test_penalty_regularisation_is_C1():
    rb = 0.04
    gb = rb / 4
    # p(0⁻) = p(0⁺), p'(0⁻) = p'(0⁺) = -1, p(ḡ) = 0 = p'(ḡ)
    lo, hi = O.regularization(-1e-14, rb, 1.0), O.regularization(1e-14, rb, 1.0)
    assert lo[0] == pytest.approx(hi[0], abs=1e-12)
    assert lo[1] == pytest.approx(hi[1], abs=1e-10) and lo[1] == -1.0
    at = O.regularization(gb, rb, 1.0)
    assert at[0] == pytest.approx(0.0, abs=1e-15) and at[1] == pytest.approx(0.0, abs=1e-12)
    # derivative of p by finite differences inside the quadratic zone
    g = 0.3 * gb
    fd = (O.regularization(g + 1e-9, rb, 1)[0] - O.regularization(g - 1e-9, rb, 1)[0]) / 2e-9
    assert fd == pytest.approx(O.regularization(g, rb, 1)[1], rel=1e-6)
    # smoothstep: η(0⁺) = ηₙ, η(ḡ) = 0, η' = dη/dg
    assert O.regularization(1e-15, rb, 2.0)[3] == pytest.approx(2.0, rel=1e-9)
    assert O.regularization(gb, rb, 2.0)[3] == pytest.approx(0.0, abs=1e-15)
    fd = (O.regularization(g + 1e-9, rb, 2.0)[3] - O.regularization(g - 1e-9, rb, 2.0)[3]) / 2e-9
    assert fd == pytest.approx(O.regularization(g, rb, 2.0)[4], rel=1e-6)
    # g ≤ 0: unchanged ηₙ, zero derivative
    assert tuple(O.regularization(-0.01, rb, 2.0)[3:]) == (2.0, 0.0)


def test_sphere_gap_is_analytic():
    eng = _engine(SphereSurface((0.1, -0.2, 0.3), 0.5), InteractionProperties(1.0, 0.0, 1.0, 0.0))
    rng = np.random.default_rng(0)
    for _ in range(10):
        p = rng.standard_normal(3)
        inside, g, n, H = eng.contact_gap(p, 0.05)
        a = p - np.array([0.1, -0.2, 0.3])
        d = np.linalg.norm(a)
        assert g == pytest.approx(d - 0.5 - 0.05, rel=1e-14)
        np.testing.assert_allclose(n, a / d, atol=1e-15)
        np.testing.assert_allclose(H, (np.eye(3) - np.outer(a, a) / d**2) / d, atol=1e-14)
        assert inside == (g <= 0.05 / 4)


def test_sdf_of_sphere_approximates_the_sphere():
    """φ = |x| - R sampled on a grid: the B-spline gap and normal approach the analytic ones."""
    sdf = DiscreteSignedDistanceField.from_function(lambda x, y, z: np.sqrt(x * x + y * y + z * z) - 0.5,
                                                    (61, 61, 61), (0.02,) * 3, (-0.6,) * 3, dtype=np.float64)
    eng = _engine(sdf, InteractionProperties(1.0, 0.0, 1.0, 0.0))
    rng = np.random.default_rng(1)
    for _ in range(20):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        p = d * rng.uniform(0.45, 0.55)
        inside, g, n, H = eng.contact_gap(p, 0.01)
        assert g == pytest.approx(np.linalg.norm(p) - 0.5 - 0.01, abs=2e-4)
        np.testing.assert_allclose(n, d, atol=5e-3)
    # outside the grid domain nothing is in contact
    assert eng.contact_gap(np.array([0.7, 0.0, 0.0]), 0.01)[0] is False


class ContactElem:
    def __init__(self, seed, friction=0.3, eta=0.0, dR=False):
        rng = np.random.default_rng(seed)
        self.X1 = np.array([0.0, 0.0, 0.0])
        self.X2 = np.array([1.0, 0.1, 0.0])
        self.l0 = float(np.linalg.norm(self.X2 - self.X1))
        self.Re0 = get_Re0(self.X1, self.X2)
        self.u1 = np.array([0.0, 0.0, -0.012]) + 0.002 * rng.standard_normal(3)
        self.u2 = np.array([0.0, 0.0, -0.006]) + 0.002 * rng.standard_normal(3)
        self.R1, self.R2 = expm(0.05 * rng.standard_normal(3)), expm(0.05 * rng.standard_normal(3))
        self.ud1, self.ud2 = 0.3 * rng.standard_normal(3), 0.3 * rng.standard_normal(3)
        self.wd1, self.wd2 = 0.3 * rng.standard_normal(3), 0.3 * rng.standard_normal(3)
        self.dR1 = expm(0.02 * rng.standard_normal(3)) if dR else np.eye(3)
        self.dR2 = expm(0.02 * rng.standard_normal(3)) if dR else np.eye(3)
        # plane z = 0 as a sphere of huge radius would be ill-conditioned: use a sphere below
        self.eng = _engine(SphereSurface((0.5, 0.0, -2.0), 2.0), InteractionProperties(50.0, friction, 0.1, eta))

    def eval(self):
        z = np.zeros(3)
        e = O.pack_elem(self.X1, self.X2, self.u1, self.u2, self.ud1, self.ud2, z, z, self.wd1, self.wd2, z, z,
                        self.R1, self.R2, self.dR1, self.dR2, self.l0, self.Re0)
        return self.eng.contact_beam(e, 0.05)

    def copy(self):
        c = ContactElem.__new__(ContactElem)
        c.__dict__ = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()}
        return c


def _perturb_config(el, k, h):
    c = el.copy()
    dv = np.zeros(3)
    dv[k % 3] = h
    blk = k // 3
    if blk == 0:
        c.u1 = c.u1 + dv
    elif blk == 1:
        c.R1 = expm(dv) @ c.R1
    elif blk == 2:
        c.u2 = c.u2 + dv
    else:
        c.R2 = expm(dv) @ c.R2
    return c


def _perturb_rate(el, k, h):
    c = el.copy()
    dv = np.zeros(3)
    dv[k % 3] = h
    name = ["ud1", "wd1", "ud2", "wd2"][k // 3]
    setattr(c, name, getattr(c, name) + dv)
    return c


def _fd_tangents(el, h=1e-7):
    Kfd = np.array([(_perturb_config(el, k, h).eval()["Tc"] - _perturb_config(el, k, -h).eval()["Tc"]) / (2 * h)
                    for k in range(12)]).T
    Cfd = np.array([(_perturb_rate(el, k, h).eval()["Tc"] - _perturb_rate(el, k, -h).eval()["Tc"]) / (2 * h)
                    for k in range(12)]).T
    return Kfd, Cfd


def _straighten(el, bend=0.0):
    """Rigidly rotated element (nodal triads follow the chord), optionally bent at node 1."""
    el.u1 = np.array([0.0, 0.0, -0.012])
    Q = expm(np.array([0.0, -0.006, 0.0]))
    el.u2 = (el.X1 + el.u1 + Q @ (el.X2 - el.X1)) - el.X2
    el.R1 = expm(bend * np.array([0.04, 0.0, 0.03])) @ Q
    el.R2 = Q.copy()
    return el


@pytest.mark.parametrize("seed,friction,eta", [(0, 0.0, 0.0), (1, 0.3, 0.0), (2, 0.3, 0.05)])
def test_contact_tangents_are_force_derivatives(seed, friction, eta):
    """Cᶜ = ∂Tᶜ/∂ḋ exactly; Kᶜ = ∂Tᶜ/∂d exactly for an unbent element. The reference's Kᶜ
    omits the variation of the Gauss point's bending offset uᵗ(Θ̅), so for a bent element the
    mismatch is proportional to the bending (halves when Θ̅ halves) — a dropped or mis-signed
    term would instead leave an O(1) mismatch on the straight element."""
    el = _straighten(ContactElem(seed, friction, eta))
    out = el.eval()
    assert out["active"] == 3
    Kfd, Cfd = _fd_tangents(el)
    assert np.linalg.norm(Kfd - out["K"]) <= 1e-7 * np.linalg.norm(out["K"])
    if friction > 0 or eta > 0:
        assert np.linalg.norm(Cfd - out["C"]) <= 1e-7 * np.linalg.norm(out["C"])
    else:
        assert np.linalg.norm(out["C"]) == 0.0
    errs = []
    for bend in (1.0, 0.5):
        el = _straighten(ContactElem(seed, friction, eta), bend)
        out = el.eval()
        Kfd, Cfd = _fd_tangents(el)
        errs.append(np.linalg.norm(Kfd - out["K"]) / np.linalg.norm(out["K"]))
        if friction > 0 or eta > 0:
            assert np.linalg.norm(Cfd - out["C"]) <= 1e-7 * np.linalg.norm(out["C"])
    assert errs[0] < 1e-2
    assert errs[0] / errs[1] == pytest.approx(2.0, rel=0.05)


def test_frictionless_contact_force_is_normal_penalty():
    el = ContactElem(3, friction=0.0, eta=0.0)
    out = el.eval()
    # total contact force on the beam = Σ over Gauss points of ω·l₀/2·kₙ pₙ n, all along n ≈ +z here
    f = out["Tc"][0:3] + out["Tc"][6:9]
    assert f[2] > 0
    assert abs(f[0]) < 0.1 * f[2] and abs(f[1]) < 0.1 * f[2]
