"""Workload builders: the reference's own test setups and the five benchmark structures.

Each builder returns (conf, params, inter) ready for run_simulation / an Engine. Values that the
workload descriptions leave open are fixed here, once, and documented in the docstrings.
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from .model import (Beams, BeamElementSurface, BeamsConfiguration, BoundaryConditions, ConcentratedForce,
                    DiscreteSignedDistanceField, Encastre, ImposedDisplacement, InteractionProperties, Loads,
                    NodesBeams, RigidInteraction, SimulationParams, SphereSurface)

GRAVITY = 9.81


def _chain(n):
    return np.array([(i, i + 1) for i in range(1, n)], dtype=np.int64)


def lumped_gravity(nodes: NodesBeams, conn: np.ndarray, rho, radius, g=GRAVITY, axis=2) -> ConcentratedForce:
    """Gravity as nodal ConcentratedForce loads (the reference's Loads has no gravity term,
    src/def_bcs_and_loads.jl:77-81): each node carries -ρA·g times half of its beams' length."""
    X = nodes.X0
    share = np.zeros(len(nodes))
    rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), (conn.shape[0],))
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (conn.shape[0],))
    for e, (a, b) in enumerate(conn):
        w = rho[e] * math.pi * radius[e] ** 2 * np.linalg.norm(X[a - 1] - X[b - 1]) / 2
        share[a - 1] += w
        share[b - 1] += w
    dofs = 6 * np.arange(len(nodes), dtype=np.int64) + axis + 1
    vals = -g * share
    return ConcentratedForce(lambda t, d, v=vals: v, dofs, vectorized=True)


# ---------------------------------------------------------------------------------------------
# reference test setups
# ---------------------------------------------------------------------------------------------

def reference_test_beams(simulation_end_time=0.15):
    """test/test_beams.jl:12-104: L-frame of 8 beams (arms of length 10 along +y then -x),
    r = 0.3, E = 1e6, ν = 0.3, ρ = 1, F_z = 100 t at dof 27, node 1 clamped, α = -0.2,
    Δt = 1e-4 fixed, tol 1e-3, max_it 20. Known answers at :113-119."""
    positions = np.array([[0.0, 0.0, 0.0], [0.0, 2.5, 0.0], [0.0, 5.0, 0.0], [0.0, 7.5, 0.0], [0.0, 10.0, 0.0],
                          [-2.5, 10.0, 0.0], [-5.0, 10.0, 0.0], [-7.5, 10.0, 0.0], [-10.0, 10.0, 0.0]])
    conn = _chain(9)
    nodes = NodesBeams(positions, plane="xy")
    beams = Beams(nodes, conn, 1e6, 0.3, 1.0, 0.3, 0.0)
    cf = ConcentratedForce(lambda t, i: 100 * t, [27])
    ndofs = 6 * len(nodes)
    bcs = BoundaryConditions(ndofs, encastre=Encastre(range(1, 7)))
    conf = BeamsConfiguration(nodes, beams, Loads(cf), bcs)
    params = SimulationParams.hht(-0.2, initial_timestep=1e-4, min_timestep=1e-4, max_timestep=1e-4,
                                  output_timestep=1e-2, simulation_end_time=simulation_end_time,
                                  tolerance_residual=1e-3, tolerance_displacement=1e-3, max_iterations=20)
    return conf, params, None


def reference_test_contact_2(simulation_end_time=2.5):
    """test/test_beams2rigid_contact_2_analytical.jl:12-128: straight wire along z (21 nodes,
    L = 0.1), dofs 2:6 clamped, dof 1 driven by a triangle wave, frictionless penalty contact
    with SphereSurface((0.05, 0.005, 0.08), 0.04), kₙ = 10, ηₙ = 1e-3, α = -0.05, Δt = 1e-2,
    tol 1e-5, max_it 5."""
    L = 10 / 100
    dx = L / 20
    n = int(round(L / dx)) + 1
    positions = np.column_stack([np.zeros(n), np.zeros(n), np.arange(n) * dx])
    conn = _chain(n)
    nodes = NodesBeams(positions, plane="xy")
    beams = Beams(nodes, conn, 5e7, 0.33, 7850.0, 3e-4, 0.0)
    maxd, tt = 10 / 100, 2.5
    v = maxd / tt

    def disp(t):
        return ((t <= tt) * v * t + (tt < t <= 2 * tt) * (-v * t + 2 * maxd)
                + (2 * tt < t <= 3 * tt) * v * (t - 2 * tt) + (3 * tt < t <= 4 * tt) * (-v * t + 4 * maxd))

    ndofs = 6 * n
    bcs = BoundaryConditions(ndofs, encastre=Encastre(range(2, 7)), imposed=ImposedDisplacement([1], disp))
    conf = BeamsConfiguration(nodes, beams, None, bcs)
    props = InteractionProperties(10.0, 0.0, 100.0, 1e-3, 10.0, 1e-3, 0.0)
    inter = RigidInteraction(SphereSurface((0.05, 0.005, 0.08), 0.04), BeamElementSurface.all(len(beams)), props)
    params = SimulationParams.hht(-0.05, initial_timestep=1e-2, min_timestep=1e-6, max_timestep=1e-2,
                                  output_timestep=1e-2, simulation_end_time=simulation_end_time,
                                  tolerance_residual=1e-5, tolerance_displacement=1e-5, max_iterations=5)
    return conf, params, inter


def reference_net(positions, connectivity, sphere=True, sdf_n=None, simulation_end_time=100.0):
    """test/test_beams2rigid_contact_1_*.jl: the net of test/input/pos_net.txt + conn_net.txt,
    initial v_y = -0.001, E = 10, ν = 0, ρ = 1500, r = 2e-4, damping 1e-3, contact kₙ = 1e-6,
    μ = 0.01, εᵗ = 0.5, ηₙ = 1e-3 against SphereSurface(0, 0.055) (analytical) or, when
    sdf_n is given, an SDF of the same sphere sampled on an sdf_n³ grid (the reference's
    sdf_sphere_20.vtk is not in its repository)."""
    positions = np.asarray(positions, dtype=np.float64)
    conn = np.asarray(connectivity, dtype=np.int64)
    v0 = np.zeros_like(positions)
    v0[:, 1] = -0.001
    nodes = NodesBeams(positions, udot0=v0, plane="xy")
    beams = Beams(nodes, conn, 10.0, 0.0, 1500.0, 2e-4, 1e-3)
    conf = BeamsConfiguration(nodes, beams, None, None)
    props = InteractionProperties(1e-6, 0.01, 0.5, 1e-3, 1e-6, 1e-3, 0.0)
    if sdf_n is None:
        master = SphereSurface((0.0, 0.0, 0.0), 0.055)
    else:
        lo, hi = -0.13, 0.13
        h = (hi - lo) / (sdf_n - 1)
        master = DiscreteSignedDistanceField.from_function(
            lambda x, y, z: np.sqrt(x * x + y * y + z * z) - 0.055, (sdf_n,) * 3, (h,) * 3, (lo,) * 3, dtype=np.float64)
    inter = RigidInteraction(master, BeamElementSurface.all(len(beams)), props)
    params = SimulationParams.hht(-0.05, initial_timestep=1e-2, max_timestep=1e-2, simulation_end_time=simulation_end_time)
    return conf, params, inter


# ---------------------------------------------------------------------------------------------
# benchmark structures
# ---------------------------------------------------------------------------------------------

def l_cantilever(tol=1e-8, load=50.0):
    """Config 1: L-shaped cantilever, two arms of length 10 (along x, then y), 10 beams per arm
    (21 nodes), r = 0.1, E = 1e6, ν = 0.3, ρ = 1, node 1 clamped, F_z at the tip = 50 t for
    t <= 1 and 0 after, α = -0.05, Δt = 0.25 fixed, t_end = 30, tolerances 1e-8. `load` replaces
    the peak tip force 50 (variants for training the learned warm start)."""
    pts = [(float(i), 0.0, 0.0) for i in range(11)] + [(10.0, float(j), 0.0) for j in range(1, 11)]
    positions = np.array(pts)
    conn = _chain(21)
    nodes = NodesBeams(positions)
    beams = Beams(nodes, conn, 1e6, 0.3, 1.0, 0.1, 0.0)
    tip_z = 6 * 20 + 3
    cf = ConcentratedForce(lambda t, i: load * t if t <= 1.0 + 1e-12 else 0.0, [tip_z])
    bcs = BoundaryConditions(126, encastre=Encastre(range(1, 7)))
    conf = BeamsConfiguration(nodes, beams, Loads(cf), bcs)
    params = SimulationParams.hht(-0.05, initial_timestep=0.25, min_timestep=1e-6, max_timestep=0.25,
                                  simulation_end_time=30.0, tolerance_residual=tol, tolerance_displacement=tol,
                                  max_iterations=20)
    return conf, params, None


def ring_drop(n_el=60, sdf_dtype=np.float32, tol=1e-8, height=0.5):
    """Config 2: ring of radius 1 (60 beams, closed), wire radius 0.05, E = 1e5, ν = 0.3,
    ρ = 1000, gravity 9.81 as nodal loads, tilted 30° about x with its lowest point at z = 0.5
    (`height`: a lower start reaches the plane within the first steps),
    falling onto the plane z = 0 given as φ = z sampled (FP32) on a 64×64×32 grid of spacing
    0.05 with origin (-1.575, -1.575, -0.25). kₙ = 1e4, μ = 0.3, εᵗ = 10, ηₙ = 0, α = -0.05,
    Δt = 1e-3, 2000 steps, no boundary conditions."""
    th = 2 * math.pi * np.arange(n_el) / n_el
    P = np.column_stack([np.cos(th), np.sin(th), np.zeros(n_el)])
    a = math.radians(30.0)
    Rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    P = P @ Rx.T
    P[:, 2] += height - P[:, 2].min()
    conn = np.array([(i, i % n_el + 1) for i in range(1, n_el + 1)], dtype=np.int64)
    nodes = NodesBeams(P)
    rho, r = 1000.0, 0.05
    beams = Beams(nodes, conn, 1e5, 0.3, rho, r, 0.0)
    loads = Loads(lumped_gravity(nodes, conn, rho, r))
    conf = BeamsConfiguration(nodes, beams, loads, None)
    sdf = DiscreteSignedDistanceField.from_function(lambda x, y, z: z, (64, 64, 32), (0.05,) * 3,
                                                    (-1.575, -1.575, -0.25), dtype=sdf_dtype)
    inter = RigidInteraction(sdf, BeamElementSurface.all(len(beams)), InteractionProperties(1e4, 0.3, 10.0, 0.0))
    params = SimulationParams.hht(-0.05, initial_timestep=1e-3, min_timestep=1e-8, max_timestep=1e-3,
                                  simulation_end_time=2.0, tolerance_residual=tol, tolerance_displacement=tol)
    return conf, params, inter


def net_on_sphere(n=40, sdf_n=128, sdf_dtype=np.float32, tol=1e-8, gap=0.1):
    """Config 3: square net of n×n wire crossings at spacing 0.05 (2n(n-1) beams), r = 0.005,
    E = 1e7, ν = 0.3, ρ = 1000, gravity as nodal loads, centred `gap` = 0.1 above the top of a rigid
    sphere of radius 0.5 at the origin, given as φ = |x| - 0.5 sampled (FP32) on sdf_n³ over
    [-1.1, 1.1]² × [-0.6, 1.6]. μ = 0.2, kₙ = 1e5, εᵗ = 10, ηₙ = 0, Δt = 5e-4, 4000 steps."""
    h = 0.05
    xs = (np.arange(n) - (n - 1) / 2) * h
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    P = np.column_stack([X.reshape(-1), Y.reshape(-1), np.full(n * n, 0.5 + gap)])
    idx = lambda i, j: j * n + i + 1
    conn = [(idx(i, j), idx(i + 1, j)) for j in range(n) for i in range(n - 1)]
    conn += [(idx(i, j), idx(i, j + 1)) for j in range(n - 1) for i in range(n)]
    conn = np.array(conn, dtype=np.int64)
    nodes = NodesBeams(P)
    rho, r = 1000.0, 0.005
    beams = Beams(nodes, conn, 1e7, 0.3, rho, r, 0.0)
    loads = Loads(lumped_gravity(nodes, conn, rho, r))
    conf = BeamsConfiguration(nodes, beams, loads, None)
    lo = np.array([-1.1, -1.1, -0.6])
    sp = (2.2 / (sdf_n - 1),) * 3
    sdf = DiscreteSignedDistanceField.from_function(lambda x, y, z: np.sqrt(x * x + y * y + z * z) - 0.5,
                                                    (sdf_n,) * 3, sp, tuple(lo), dtype=sdf_dtype)
    inter = RigidInteraction(sdf, BeamElementSurface.all(len(beams)), InteractionProperties(1e5, 0.2, 10.0, 0.0))
    params = SimulationParams.hht(-0.05, initial_timestep=5e-4, min_timestep=1e-9, max_timestep=5e-4,
                                  simulation_end_time=2.0, tolerance_residual=tol, tolerance_displacement=tol)
    return conf, params, inter


def _helix_frames(R, pitch, phase, handed, s):
    """Points and [tangent, inward normal, binormal] frames of a helix at parameters s (turn angle)."""
    c = pitch / (2 * math.pi)
    ang = phase + handed * s
    P = np.column_stack([R * np.cos(ang), R * np.sin(ang), c * s])
    T = np.column_stack([-handed * R * np.sin(ang), handed * R * np.cos(ang), np.full_like(s, c)])
    T /= np.linalg.norm(T, axis=1)[:, None]
    Nn = np.column_stack([-np.cos(ang), -np.sin(ang), np.zeros_like(s)])
    B = np.cross(T, Nn)
    F = np.stack([T, Nn, B], axis=2)  # columns
    return P, F


def braided_stent(n_wires=24, radius=4e-3, pitch=8e-3, length=40e-3, el_per_wire=200, crimp_radius=1.5e-3,
                  vessel_radius=3e-3, sdf_n=256, sdf_dtype=np.float32, wire_radius=5e-5, E=50e9, rho=6500.0,
                  mu=0.1, kn=None, dt=1e-3, tol=1e-8, seed=None):
    """Config 4: braided stent of n_wires helices (half right-, half left-handed, phases 2πk/(n/2)),
    radius 4 mm, pitch 8 mm, length 40 mm, 200 beams per wire, wire radius 0.05 mm, E = 50 GPa,
    ν = 0.3, ρ = 6500. The stress-free helix is the reference configuration X₀; the initial state
    is the same wire crimped to 1.5 mm radius at constant wire length (pitch lengthened so each
    turn keeps its arc length), with nodal rotations R⁰ = F_c F₀ᵀ taking the stress-free helix
    frame onto the crimped one (the reference has no crimping procedure; NodesBeams accepts u⁰
    and R, src/beams/def_node_beam.jl:55-74). Released inside a rigid vessel φ = 3 mm - ρ_xy
    sampled (FP32) on sdf_n³ covering the crimped length. μ = 0.1, εᵗ = 10 (m/s)², ηₙ = 0,
    kₙ = 1e3 N/m² unless given, Δt = 1e-3. Wires interact only through the vessel."""
    rng = np.random.default_rng(seed) if seed is not None else None
    if rng is not None:  # seeded geometry variants: R ±5 %, pitch ±10 %
        radius *= 1 + 0.05 * (2 * rng.random() - 1)
        pitch *= 1 + 0.10 * (2 * rng.random() - 1)
    turns = length / pitch
    s = np.linspace(0.0, 2 * math.pi * turns, el_per_wire + 1)
    Lturn = math.hypot(2 * math.pi * radius, pitch)
    pitch_c = math.sqrt(max(Lturn**2 - (2 * math.pi * crimp_radius) ** 2, (0.1 * pitch) ** 2))
    half = n_wires // 2
    Xs, Us, Rs, conn = [], [], [], []
    off = 0
    for k in range(n_wires):
        handed = 1.0 if k < half else -1.0
        phase = 2 * math.pi * (k % half) / half
        P0, F0 = _helix_frames(radius, pitch, phase, handed, s)
        Pc, Fc = _helix_frames(crimp_radius, pitch_c, phase, handed, s)
        Xs.append(P0)
        Us.append(Pc - P0)
        Rs.append(np.einsum("nij,nkj->nik", Fc, F0))
        conn.extend((off + i + 1, off + i + 2) for i in range(el_per_wire))
        off += el_per_wire + 1
    X = np.vstack(Xs)
    U = np.vstack(Us)
    R0 = np.concatenate(Rs, axis=0)
    conn = np.array(conn, dtype=np.int64)
    nodes = NodesBeams(X, u0=U, R=R0)
    beams = Beams(nodes, conn, E, 0.3, rho, wire_radius, 0.0)
    conf = BeamsConfiguration(nodes, beams, None, None)
    zc = X[:, 2] + U[:, 2]
    zlo, zhi = zc.min() - 2e-3, zc.max() + 2e-3
    xy = vessel_radius + 0.5e-3
    sp = (2 * xy / (sdf_n - 1), 2 * xy / (sdf_n - 1), (zhi - zlo) / (sdf_n - 1))
    sdf = DiscreteSignedDistanceField.from_function(lambda x, y, z: vessel_radius - np.sqrt(x * x + y * y),
                                                    (sdf_n,) * 3, sp, (-xy, -xy, zlo), dtype=sdf_dtype)
    kn = 1e3 if kn is None else kn
    inter = RigidInteraction(sdf, BeamElementSurface.all(len(beams)), InteractionProperties(kn, mu, 10.0, 0.0))
    params = SimulationParams.hht(-0.05, initial_timestep=dt, min_timestep=1e-9, max_timestep=dt,
                                  simulation_end_time=2.0, tolerance_residual=tol, tolerance_displacement=tol)
    return conf, params, inter


def stent_sweep_instance(i, sdf_n=128, length=40e-3):
    """Config 5, instance i: a braided stent drawn with numpy's PCG64 seeded 1000 + i: wire count
    W ~ U{12, 16, ..., 48}, radius R ~ U[3, 5] mm, pitch ~ U[6, 10] mm, beams per wire
    ~ U{100, ..., 300} (1 200 to 14 400 beams), in its own cylindrical vessel of radius 0.7 R
    (FP32 SDF on sdf_n³). The stent is crimped to 0.375 R (config 4's 1.5 mm of 4 mm) and released;
    length 40 mm and every material, contact and time-integration value as config 4."""
    rng = np.random.default_rng(1000 + i)
    W = int(rng.choice(np.arange(12, 49, 4)))
    R = float(rng.uniform(3e-3, 5e-3))
    pitch = float(rng.uniform(6e-3, 10e-3))
    m = int(rng.integers(100, 301))
    return braided_stent(n_wires=W, radius=R, pitch=pitch, length=length, el_per_wire=m, crimp_radius=0.375 * R,
                         vessel_radius=0.7 * R, sdf_n=sdf_n)


def stent_sweep(n_instances=512, sdf_n=128, first=0):
    """Config 5: instances first .. first + n_instances - 1 of the design sweep."""
    return [stent_sweep_instance(first + i, sdf_n=sdf_n) for i in range(n_instances)]


WORKLOADS = {
    "l_cantilever": l_cantilever,
    "ring_drop": ring_drop,
    "net_on_sphere": net_on_sphere,
    "braided_stent": braided_stent,
    "stent_sweep": stent_sweep,
}
