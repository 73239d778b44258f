"""Model construction mirroring EndoBeams.jl's user-facing constructors.

Every class here restates a reference constructor (cited per class) and produces plain numpy
arrays; `model_desc` packs them into the C ABI's eb_model_desc. Nothing in this module touches a
device.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import abi

EPS = np.finfo(np.float64).eps


# ---------------------------------------------------------------------------------------------
# nodes and beams
# ---------------------------------------------------------------------------------------------

class NodesBeams:
    """NodesBeams(X, u⁰, u̇⁰, ü⁰, w⁰, ẇ⁰, ẅ⁰, plane=nothing, R=nothing)
    (src/beams/def_node_beam.jl:55-128). Arrays are (N, 3); R is (N, 3, 3) or None (identity).
    Node i owns global dofs 6(i-1)+1 … 6(i-1)+6 (1-based, :62-67)."""

    def __init__(self, X, u0=None, udot0=None, uddot0=None, w0=None, wdot0=None, wddot0=None, plane=None, R=None):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        n = X.shape[0]
        z = lambda a: np.zeros((n, 3)) if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(n, 3)
        self.X0 = X
        self.u, self.udot, self.uddot = z(u0), z(udot0), z(uddot0)
        self.w, self.wdot, self.wddot = z(w0), z(wdot0), z(wddot0)
        if R is None:
            self.R = np.tile(np.eye(3), (n, 1, 1))
        else:
            self.R = np.ascontiguousarray(R, dtype=np.float64).reshape(n, 3, 3)
        self.plane = plane

    def __len__(self):
        return self.X0.shape[0]


def get_Re0(X1, X2):
    """get_Rₑ⁰ (src/beams/def_beam.jl:93-116): rotation taking e₁ onto the chord direction."""
    X1 = np.asarray(X1, dtype=np.float64)
    X2 = np.asarray(X2, dtype=np.float64)
    l0 = np.linalg.norm(X2 - X1)
    E1 = np.array([1.0, 0.0, 0.0])
    E10 = (X2 - X1) / l0
    v = np.cross(E1, E10)
    s = np.linalg.norm(v)
    c = float(E1 @ E10)
    if c < -(1 - 2 * EPS):
        return np.diag([-1.0, 1.0, -1.0])
    if c > (1 - 2 * EPS):
        return np.eye(3)
    Sv = skew(v)
    return (np.eye(3) + Sv) + ((1 - c) / s**2 * Sv) @ Sv


def get_Re0_many(X1, X2):
    """get_Re0 for rows of chord endpoints at once (the same formula per row)."""
    X1 = np.asarray(X1, dtype=np.float64).reshape(-1, 3)
    X2 = np.asarray(X2, dtype=np.float64).reshape(-1, 3)
    d = X2 - X1
    l0 = np.sqrt(np.einsum("ij,ij->i", d, d))
    E10 = d / l0[:, None]
    # v = e₁ × E10, c = e₁ · E10
    v = np.column_stack([np.zeros_like(l0), -E10[:, 2], E10[:, 1]])
    s2 = v[:, 1] ** 2 + v[:, 2] ** 2
    c = E10[:, 0]
    Sv = np.zeros((len(l0), 3, 3))
    Sv[:, 0, 1], Sv[:, 0, 2] = -v[:, 2], v[:, 1]
    Sv[:, 1, 0], Sv[:, 1, 2] = v[:, 2], -v[:, 0]
    Sv[:, 2, 0], Sv[:, 2, 1] = -v[:, 1], v[:, 0]
    out = np.eye(3) + Sv + ((1 - c) / np.where(s2 > 0, s2, 1.0))[:, None, None] * (Sv @ Sv)
    out[c < -(1 - 2 * EPS)] = np.diag([-1.0, 1.0, -1.0])
    out[c > (1 - 2 * EPS)] = np.eye(3)
    return out


def skew(v):
    return np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])


def beam_material_properties(l0, E, nu, rho, radius, damping):
    """BeamMaterialProperties + K̄ⁱⁿᵗ_beam (src/beams/def_beam.jl:74-85, :135-149) packed as the
    14-double record of the ABI: radius, E, K̄ū, K̄Θ̅(3), K̄Θ̅Θ̅(3), Jᵨ(3), Aᵨ, damping."""
    G = E / (2 * (1 + nu))
    A = math.pi * radius**2
    I22 = math.pi * radius**4 / 4
    I33 = I22
    Io = I22 + I33
    Ku = A * E / l0
    KT = (G * Io / l0, 4 * E * I33 / l0, 4 * E * I22 / l0)
    KTT = (-G * Io / l0, 2 * E * I33 / l0, 2 * E * I22 / l0)
    J = (rho * Io, rho * I22, rho * I33)
    return np.array([radius, E, Ku, *KT, *KTT, *J, rho * A, damping], dtype=np.float64)


def beam_material_properties_many(l0, E, nu, rho, radius, damping):
    """beam_material_properties for arrays of beams (rows of the same 14-double record)."""
    G = E / (2 * (1 + nu))
    A = math.pi * radius**2
    I22 = math.pi * radius**4 / 4
    I33 = I22
    Io = I22 + I33
    return np.column_stack([radius, E, A * E / l0, G * Io / l0, 4 * E * I33 / l0, 4 * E * I22 / l0,
                            -G * Io / l0, 2 * E * I33 / l0, 2 * E * I22 / l0, rho * Io, rho * I22, rho * I33,
                            rho * A, damping]).astype(np.float64)


class Beams:
    """Beams(nodes, connectivity, E, ν, ρ, radius, damping, Rₑ⁰=nothing)
    (src/beams/def_beam.jl:32-72). Scalars or per-beam arrays; connectivity is (E, 2), 1-based."""

    def __init__(self, nodes: NodesBeams, connectivity, E, nu, rho, radius, damping, Re0=None):
        conn = np.ascontiguousarray(connectivity, dtype=np.int64).reshape(-1, 2)
        ne = conn.shape[0]
        per = lambda a: np.broadcast_to(np.asarray(a, dtype=np.float64), (ne,))
        E, nu, rho, radius, damping = map(per, (E, nu, rho, radius, damping))
        X = nodes.X0
        self.conn = conn
        X1, X2 = X[conn[:, 0] - 1], X[conn[:, 1] - 1]
        d = X1 - X2
        self.l0 = np.sqrt(np.einsum("ij,ij->i", d, d))
        self.Re0 = get_Re0_many(X1, X2) if Re0 is None else np.asarray(Re0, dtype=np.float64).reshape(ne, 3, 3).copy()
        self.props = beam_material_properties_many(self.l0, E, nu, rho, radius, damping)

    def __len__(self):
        return self.conn.shape[0]


# ---------------------------------------------------------------------------------------------
# boundary conditions and loads (src/def_bcs_and_loads.jl)
# ---------------------------------------------------------------------------------------------

@dataclass
class Encastre:
    """Encastre(blocked_dofs) (src/def_bcs_and_loads.jl:2-4); dofs are 1-based."""
    blocked_dofs: Sequence[int]


@dataclass
class ImposedDisplacement:
    """ImposedDisplacement(displaced_dofs, f(t)) (src/def_bcs_and_loads.jl:7-10)."""
    displaced_dofs: Sequence[int]
    imposed_displacements_function: Callable[[float], float]


class BoundaryConditions:
    """BoundaryConditions (src/def_bcs_and_loads.jl:13-63): fixed/free dofs, the indices of the
    imposed-displacement dofs among the fixed ones and their current values."""

    def __init__(self, ndofs: int, encastre: Optional[Encastre] = None, imposed: Optional[ImposedDisplacement] = None):
        enc = [] if encastre is None else [int(d) for d in np.atleast_1d(encastre.blocked_dofs)]
        imp = [] if imposed is None else [int(d) for d in np.atleast_1d(imposed.displaced_dofs)]
        if imposed is not None and encastre is not None:
            fixed = sorted(imp + enc)
        elif imposed is not None:
            fixed = list(imp)
        else:
            fixed = list(enc)
        self.fixed_dofs = np.array(fixed, dtype=np.int64)
        fs = set(fixed)
        self.free_dofs = np.array([d for d in range(1, ndofs + 1) if d not in fs], dtype=np.int64)
        self.imposed_displacements = np.zeros(len(fixed))
        self.displaced_indices = np.array([i for i, d in enumerate(fixed) if d in set(imp)], dtype=np.int64)
        self.imposed_displacements_function = None if imposed is None else imposed.imposed_displacements_function


@dataclass
class ConcentratedForce:
    """ConcentratedForce(force_function(t, dof), loaded_dofs) (src/def_bcs_and_loads.jl:66-74).
    With vectorized=True, force_function(t, dofs_array) returns all values at once."""
    force_function: Callable
    loaded_dofs: Sequence[int]
    vectorized: bool = False

    def values(self, t: float) -> np.ndarray:
        dofs = np.asarray(self.loaded_dofs, dtype=np.int64)
        if self.vectorized:
            return np.ascontiguousarray(np.broadcast_to(self.force_function(t, dofs), dofs.shape), dtype=np.float64)
        return np.array([self.force_function(t, int(d)) for d in dofs], dtype=np.float64)


@dataclass
class Loads:
    """Loads(concentrated_force) (src/def_bcs_and_loads.jl:77-81)."""
    concentrated_force: ConcentratedForce


class BeamsConfiguration:
    """BeamsConfiguration(nodes, beams, loads, bcs) (src/def_configurations.jl:5-25)."""

    def __init__(self, nodes: NodesBeams, beams: Beams, loads: Optional[Loads] = None, bcs: Optional[BoundaryConditions] = None):
        self.nodes = nodes
        self.beams = beams
        self.ndofs = 6 * len(nodes)
        self.offset_dofs = 0
        self.loads = loads
        self.bcs = bcs if bcs is not None else BoundaryConditions(self.ndofs)


# ---------------------------------------------------------------------------------------------
# parameters (src/def_params.jl)
# ---------------------------------------------------------------------------------------------

@dataclass
class SimulationParams:
    """SimulationParams (src/def_params.jl:1-35); α, β, γ have no default, as in the reference.
    krylov_* control the device linear solve (no counterpart: the reference calls Pardiso)."""
    alpha: float
    beta: float
    gamma: float
    verbose: bool = False
    record_timings: bool = False
    output_dir: str = ""
    initial_timestep: float = 1e-2
    min_timestep: float = 1e-10
    max_timestep: float = 1e-2
    output_timestep: float = 1e-4
    simulation_end_time: float = 1e-2
    accelerate_after_success_it: int = 4
    stop_on_energy_threshold: bool = False
    energy_threshold: float = 0.0
    tcompt_max: float = 1800.0
    stop_long_simulation: bool = False
    tolerance_residual: float = 1e-5
    tolerance_displacement: float = 1e-5
    max_iterations: int = 20
    n_gauss_beams: int = 3
    w_gauss_beams: tuple = (5 / 9, 8 / 9, 5 / 9)
    z_gauss_beams: tuple = (-math.sqrt(3 / 5), 0.0, math.sqrt(3 / 5))
    krylov_tol: float = 1e-11
    krylov_maxit: int = 200

    @staticmethod
    def hht(alpha: float, **kw) -> "SimulationParams":
        """β = (1-α)²/4, γ = (1-2α)/2 as every reference script sets them (e.g. test/test_beams.jl:85-87)."""
        return SimulationParams(alpha=alpha, beta=0.25 * (1 - alpha) ** 2, gamma=0.5 * (1 - 2 * alpha), **kw)

    def newton(self) -> abi.NewtonParams:
        return abi.NewtonParams(self.alpha, self.beta, self.gamma, self.tolerance_residual,
                                self.tolerance_displacement, self.max_iterations, self.krylov_maxit, self.krylov_tol)


# ---------------------------------------------------------------------------------------------
# contact (src/interactions/def_interactions.jl)
# ---------------------------------------------------------------------------------------------

@dataclass
class InteractionProperties:
    """InteractionProperties(kₙ, μ, εᵗ, ηₙ, kₜ, ηₜ, u̇ₛ) (def_interactions.jl:6-14)."""
    kn: float
    mu: float
    eps_t: float
    eta_n: float
    kt: float = 0.0
    eta_t: float = 0.0
    us: float = 0.0


@dataclass
class SphereSurface:
    """SphereSurface(center, radius) (def_interactions.jl:30-33)."""
    center: Sequence[float]
    radius: float


class DiscreteSignedDistanceField:
    """DiscreteSignedDistanceField (def_interactions.jl:36-43, :67-82) from a regular grid of
    samples (x fastest, as read_vtk_sdf returns them). `inside=False` flips the sign (:71)."""

    def __init__(self, samples, n, spacing, origin, inside=True):
        s = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1)
        n = tuple(int(v) for v in n)
        if s.size != n[0] * n[1] * n[2]:
            raise ValueError("sample count does not match grid dimensions")
        self.samples = s if inside else -s
        self.prefiltered = False  # True: `samples` holds the padded B-spline coefficients
        self.n = n
        self.spacing = tuple(float(v) for v in spacing)
        self.origin = tuple(float(v) for v in origin)
        self.dom = tuple(v for a in range(3) for v in (self.origin[a], self.origin[a] + (n[a] - 1) * self.spacing[a]))

    @staticmethod
    def from_vtk(filename, inside=True):
        npx, npy, npz, dx, dy, dz, dom, sdf = read_vtk_sdf(filename)
        return DiscreteSignedDistanceField(sdf, (npx, npy, npz), (dx, dy, dz), (dom[0], dom[2], dom[4]), inside)

    @staticmethod
    def from_function(phi, n, spacing, origin, dtype=np.float32, inside=True):
        """Samples phi(x, y, z) on the grid, stored at `dtype` (FP32 voxel grids) before use."""
        axes = [origin[a] + spacing[a] * np.arange(n[a]) for a in range(3)]
        Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        vals = np.asarray(phi(X, Y, Z), dtype=np.float64).astype(dtype).astype(np.float64)
        return DiscreteSignedDistanceField(vals.reshape(-1), n, spacing, origin, inside)


def read_vtk_sdf(filename):
    """read_vtk_sdf (src/IO/read_input_files.jl:154-194): legacy ASCII VTK structured points."""
    with open(filename, "r") as fid:
        for _ in range(4):
            fid.readline()
        nx, ny, nz = (int(v) for v in fid.readline().split()[1:4])
        dx, dy, dz = (float(v) for v in fid.readline().split()[1:4])
        x0, y0, z0 = (float(v) for v in fid.readline().split()[1:4])
        dom = (x0, x0 + (nx - 1) * dx, y0, y0 + (ny - 1) * dy, z0, z0 + (nz - 1) * dz)
        nnode = int(fid.readline().split()[1])
        fid.readline()
        fid.readline()
        vals = []
        while len(vals) < nnode:
            vals.extend(float(v) for v in fid.readline().split())
    return nx, ny, nz, dx, dy, dz, dom, np.array(vals[:nnode])


@dataclass
class BeamElementSurface:
    """BeamElementSurface(contact_beams) (def_interactions.jl:46-56); 1-based beam indices."""
    contact_beams: Sequence[int]

    @staticmethod
    def all(nbeams: int) -> "BeamElementSurface":
        return BeamElementSurface(list(range(1, nbeams + 1)))


@dataclass
class RigidInteraction:
    """RigidInteraction(master, slave, properties) (def_interactions.jl:95-99)."""
    master: object
    slave: BeamElementSurface
    properties: InteractionProperties


# ---------------------------------------------------------------------------------------------
# ABI packing
# ---------------------------------------------------------------------------------------------

class ModelArrays:
    """Keeps the numpy buffers alive for as long as the eb_model_desc that points at them."""

    def __init__(self, conf: BeamsConfiguration, params: SimulationParams, inter: Optional[RigidInteraction] = None,
                 precision: int = abi.EB_PRECISION_FP64):
        nodes, beams = conf.nodes, conf.beams
        self.X0 = np.ascontiguousarray(nodes.X0.reshape(-1))
        self.conn = np.ascontiguousarray(beams.conn.reshape(-1))
        self.l0 = np.ascontiguousarray(beams.l0)
        # Rₑ⁰ column-major per beam
        self.Re0 = np.ascontiguousarray(np.transpose(beams.Re0, (0, 2, 1)).reshape(-1))
        self.props = np.ascontiguousarray(beams.props.reshape(-1))
        self.fixed = np.ascontiguousarray(conf.bcs.fixed_dofs, dtype=np.int64)
        self.gw = np.ascontiguousarray(params.w_gauss_beams[: params.n_gauss_beams], dtype=np.float64)
        self.gz = np.ascontiguousarray(params.z_gauss_beams[: params.n_gauss_beams], dtype=np.float64)
        d = abi.ModelDesc()
        d.abi_version = abi.EB_ABI_VERSION
        d.precision = precision
        d.n_nodes = len(nodes)
        d.n_beams = len(beams)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        lp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        d.X0, d.conn, d.l0, d.Re0, d.props = dp(self.X0), lp(self.conn), dp(self.l0), dp(self.Re0), dp(self.props)
        d.n_fixed = self.fixed.size
        d.fixed_dofs = lp(self.fixed) if self.fixed.size else None
        d.n_gauss = params.n_gauss_beams
        d.gauss_w, d.gauss_z = dp(self.gw), dp(self.gz)
        d.surface = abi.EB_SURFACE_NONE
        if inter is not None:
            p = inter.properties
            d.kn, d.mu, d.eps_t, d.eta_n = p.kn, p.mu, p.eps_t, p.eta_n
            self.cbeams = np.ascontiguousarray(inter.slave.contact_beams, dtype=np.int64)
            d.n_contact_beams = self.cbeams.size
            d.contact_beams = lp(self.cbeams)
            m = inter.master
            if isinstance(m, SphereSurface):
                d.surface = abi.EB_SURFACE_SPHERE
                for k in range(3):
                    d.sphere_center[k] = float(m.center[k])
                d.sphere_radius = float(m.radius)
            elif isinstance(m, DiscreteSignedDistanceField):
                d.surface = abi.EB_SURFACE_SDF
                d.sdf_prefiltered = 1 if getattr(m, "prefiltered", False) else 0
                self.sdf = np.ascontiguousarray(m.samples, dtype=np.float64)
                for k in range(3):
                    d.sdf_n[k] = m.n[k]
                    d.sdf_d[k] = m.spacing[k]
                    d.sdf_x0[k] = m.origin[k]
                d.sdf_samples = dp(self.sdf)
            else:
                raise TypeError(f"unsupported master surface {type(m).__name__}")
        self.desc = d


def initial_fields(conf: BeamsConfiguration):
    """Node fields as NodesBeams initialises them: current and step-n copies equal, ΔR = I."""
    n = conf.nodes
    N = len(n)
    Rcm = np.ascontiguousarray(np.transpose(n.R, (0, 2, 1)).reshape(-1))  # column-major per node
    I = np.tile(np.eye(3).reshape(-1), N)
    out = {}
    for suffix in ("", "_N"):
        out["U" + suffix] = n.u.reshape(-1)
        out["UDOT" + suffix] = n.udot.reshape(-1)
        out["UDDOT" + suffix] = n.uddot.reshape(-1)
        out["W" + suffix] = n.w.reshape(-1)
        out["WDOT" + suffix] = n.wdot.reshape(-1)
        out["WDDOT" + suffix] = n.wddot.reshape(-1)
        out["R" + suffix] = Rcm
        out["DR" + suffix] = I
    return out
