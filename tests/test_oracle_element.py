"""Pins of the oracle's corotational beam element (compute_beams, src/beams/compute_beams.jl)
against what the mechanics fixes, so that a dropped term, a wrong sign or a transposed block fails:
  * Tⁱⁿᵗ is the gradient of the strain energy for spatial-spin rotation variations;
  * Kⁱⁿᵗ is the derivative of Tⁱⁿᵗ (same variations);
  * small-strain limit = the textbook 3D Euler–Bernoulli 12×12 stiffness in the element frame;
  * rigid-body invariance of the energy and objectivity of Tⁱⁿᵗ;
  * M = ∂Tᵏ/∂d̈, Cᵏ = ∂Tᵏ/∂ḋ, kinetic energy = ½ ḋᵀMḋ (ΔR = I).
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from endobeams_amd.model import beam_material_properties, get_Re0

GW = (5 / 9, 8 / 9, 5 / 9)
GZ = (-np.sqrt(3 / 5), 0.0, np.sqrt(3 / 5))


def expm(v):
    return Rotation.from_rotvec(v).as_matrix()


class Elem:
    """One beam with a random deformed configuration and random nodal rates."""

    def __init__(self, seed, rot=0.2, disp=0.05, rate=1.0, damping=0.0, l0=1.3, E=2e5, r=0.05):
        rng = np.random.default_rng(seed)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        self.X1 = rng.standard_normal(3)
        self.X2 = self.X1 + l0 * d
        self.l0 = float(np.linalg.norm(self.X2 - self.X1))
        self.Re0 = get_Re0(self.X1, self.X2)
        self.props = beam_material_properties(self.l0, E, 0.3, 7.0, r, damping)
        self.u1, self.u2 = disp * rng.standard_normal(3), disp * rng.standard_normal(3)
        self.R1, self.R2 = expm(rot * rng.standard_normal(3)), expm(rot * rng.standard_normal(3))
        self.v = [rate * rng.standard_normal(3) for _ in range(8)]  # ud1 ud2 udd1 udd2 wd1 wd2 wdd1 wdd2
        self.dR1, self.dR2 = np.eye(3), np.eye(3)

    def pack(self):
        ud1, ud2, udd1, udd2, wd1, wd2, wdd1, wdd2 = self.v
        return O.pack_elem(self.X1, self.X2, self.u1, self.u2, ud1, ud2, udd1, udd2, wd1, wd2, wdd1, wdd2,
                           self.R1, self.R2, self.dR1, self.dR2, self.l0, self.Re0)

    def eval(self):
        return O.compute_beam(self.pack(), self.props, GW, GZ)

    def perturb(self, k, h):
        """Configuration variation along local dof k (u₁ θ₁ u₂ θ₂), spins applied on the left."""
        e = Elem.__new__(Elem)
        e.__dict__ = {kk: (vv.copy() if isinstance(vv, np.ndarray) else (list(vv) if isinstance(vv, list) else vv))
                      for kk, vv in self.__dict__.items()}
        dv = np.zeros(3)
        dv[k % 3] = h
        blk = k // 3
        if blk == 0:
            e.u1 = e.u1 + dv
        elif blk == 1:
            e.R1 = expm(dv) @ e.R1
        elif blk == 2:
            e.u2 = e.u2 + dv
        else:
            e.R2 = expm(dv) @ e.R2
        return e


def fd_columns(elem, fn, h):
    cols = []
    for k in range(12):
        cols.append((fn(elem.perturb(k, h)) - fn(elem.perturb(k, -h))) / (2 * h))
    return np.array(cols).T


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_internal_force_is_energy_gradient(seed):
    el = Elem(seed)
    out = el.eval()
    g = fd_columns(el, lambda e: np.array([e.eval()["strain"]]), 1e-6)[0]
    assert np.linalg.norm(g - out["Tint"]) <= 1e-6 * np.linalg.norm(out["Tint"])


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4])
def test_stiffness_is_force_derivative(seed):
    """The reference's tangent matches ∂Tⁱⁿᵗ/∂d up to a remainder of third order in the
    deformation (the averaged-triad local frame is linearised to second order): the absolute
    mismatch must fall by ≥ 4x when rotations halve, and be ≤ 2e-6 relative at 0.1 rad. A dropped
    or mis-signed tangent term is first order (halving ratio ≈ 2, relative error ≳ 1e-3)."""
    errs = []
    for rot in (0.1, 0.05):
        el = Elem(seed, rot=rot, disp=rot / 4)
        out = el.eval()
        J = fd_columns(el, lambda e: e.eval()["Tint"], 1e-6)
        errs.append(np.linalg.norm(J - out["K"]))
        if rot == 0.1:
            assert errs[-1] <= 2e-6 * np.linalg.norm(out["K"])
    assert errs[0] >= 4 * errs[1]


def test_small_strain_limit_is_euler_bernoulli():
    """Undeformed element: Kⁱⁿᵗ = Tᵀ K_EB T with T = blockdiag(Rₑ⁰ᵀ) and the textbook 3D frame
    stiffness (axial EA/L, torsion GJ/L, bending 12EI/L³, 6EI/L², 4EI/L, 2EI/L)."""
    el = Elem(5, rot=0.0, disp=0.0)
    el.u1[:] = 0
    el.u2[:] = 0
    el.R1, el.R2 = np.eye(3), np.eye(3)
    out = el.eval()
    E, nu, r, L = 2e5, 0.3, 0.05, el.l0
    G = E / (2 * (1 + nu))
    A, I = np.pi * r**2, np.pi * r**4 / 4
    J = 2 * I
    k = np.zeros((12, 12))
    k[0, 0] = k[6, 6] = E * A / L
    k[0, 6] = k[6, 0] = -E * A / L
    k[3, 3] = k[9, 9] = G * J / L
    k[3, 9] = k[9, 3] = -G * J / L
    for (v, th, s) in [(1, 5, 1.0), (2, 4, -1.0)]:  # bending in local x-y (v, θz) and x-z (w, θy)
        a, b, c, d = 12 * E * I / L**3, 6 * E * I / L**2, 4 * E * I / L, 2 * E * I / L
        idx = [v, th, v + 6, th + 6]
        m = np.array([[a, s * b, -a, s * b], [s * b, c, -s * b, d], [-a, -s * b, a, -s * b], [s * b, d, -s * b, c]])
        for i in range(4):
            for j in range(4):
                k[idx[i], idx[j]] = m[i, j]
    T = np.kron(np.eye(4), el.Re0.T)
    Kexp = T.T @ k @ T
    assert np.linalg.norm(out["K"] - Kexp) <= 1e-10 * np.linalg.norm(Kexp)
    assert np.linalg.norm(out["Tint"]) <= 1e-12 * np.linalg.norm(Kexp)


def test_rigid_motion_invariance_and_objectivity():
    el = Elem(7)
    out = el.eval()
    Q = expm(np.array([0.3, -1.1, 0.7]))
    el2 = Elem(7)
    # rotate the current configuration rigidly about the origin: x → Q x, R → Q R
    x1, x2 = Q @ (el.X1 + el.u1), Q @ (el.X2 + el.u2)
    el2.u1, el2.u2 = x1 - el.X1, x2 - el.X2
    el2.R1, el2.R2 = Q @ el.R1, Q @ el.R2
    out2 = el2.eval()
    assert out2["strain"] == pytest.approx(out["strain"], rel=1e-10)
    Qb = np.kron(np.eye(4), Q)
    assert np.linalg.norm(out2["Tint"] - Qb @ out["Tint"]) <= 1e-9 * np.linalg.norm(out["Tint"])


def _with_rates(el, ud, wd, udd, wdd):
    e = Elem.__new__(Elem)
    e.__dict__ = dict(el.__dict__)
    e.v = [ud[:3], ud[3:], udd[:3], udd[3:], wd[:3], wd[3:], wdd[:3], wdd[3:]]
    return e


def _rates(el):
    ud1, ud2, udd1, udd2, wd1, wd2, wdd1, wdd2 = el.v
    d = np.concatenate([ud1, wd1, ud2, wd2])
    dd = np.concatenate([udd1, wdd1, udd2, wdd2])
    return d, dd


def _split(v):
    return np.concatenate([v[0:3], v[6:9]]), np.concatenate([v[3:6], v[9:12]])


@pytest.mark.parametrize("seed", [0, 3])
def test_mass_is_acceleration_derivative(seed):
    el = Elem(seed)
    d, dd = _rates(el)

    def Tk(acc):
        ud, wd = _split(d)
        udd, wdd = _split(acc)
        return O.compute_beam(_with_rates(el, ud, wd, udd, wdd).pack(), el.props, GW, GZ)["Tk"]

    Mfd = np.array([(Tk(dd + 1e-3 * np.eye(12)[k]) - Tk(dd - 1e-3 * np.eye(12)[k])) / 2e-3 for k in range(12)]).T
    M = el.eval()["M"]
    assert np.linalg.norm(Mfd - M) <= 1e-9 * np.linalg.norm(M)


@pytest.mark.parametrize("seed,damping", [(0, 0.0), (4, 0.0), (6, 0.7)])
def test_gyroscopic_matrix_is_velocity_derivative(seed, damping):
    el = Elem(seed, damping=damping)
    d, dd = _rates(el)

    def Tk(vel):
        ud, wd = _split(vel)
        udd, wdd = _split(dd)
        return O.compute_beam(_with_rates(el, ud, wd, udd, wdd).pack(), el.props, GW, GZ)["Tk"]

    h = 1e-6
    Cfd = np.array([(Tk(d + h * np.eye(12)[k]) - Tk(d - h * np.eye(12)[k])) / (2 * h) for k in range(12)]).T
    Cm = el.eval()["C"]
    assert np.linalg.norm(Cfd - Cm) <= 1e-6 * np.linalg.norm(Cm)


def test_kinetic_energy_is_quadratic_form_of_mass():
    el = Elem(9)
    out = el.eval()
    d, _ = _rates(el)
    assert out["kinetic"] == pytest.approx(0.5 * d @ out["M"] @ d, rel=1e-12)
