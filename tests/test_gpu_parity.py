"""Parity of libendobeams_hip.so (through its C ABI) with the CPU oracle.

Every reference function on the hot path is called on both engines from the same randomised
state and its outputs compared norm-wise (‖a-b‖/‖b‖): 1e-10 for the FP64 kernels (bar: north
star), 1e-8 on the linear-solve output (BiCGStab to 1e-11 vs a direct LU). Trajectories are
compared after many steps with identical Newton iteration counts.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import os

import oracle as O
from conftest import REPO, rel_err
from endobeams_amd import HipEngine, configs, run_simulation

pytestmark = pytest.mark.gpu

TOL = 1e-10
GOLDEN = os.path.join(REPO, "tests", "golden")


def small_stent():
    return configs.braided_stent(n_wires=4, el_per_wire=40, sdf_n=64)


def clamped_stent():
    """Stent wires with both ends clamped and an imposed displacement: exercises the BC lift."""
    from endobeams_amd.model import BoundaryConditions, Encastre, ImposedDisplacement, BeamsConfiguration
    conf, params, inter = configs.braided_stent(n_wires=2, el_per_wire=30, sdf_n=48, crimp_radius=4e-3)
    nd = conf.ndofs
    enc = Encastre(list(range(2, 7)) + list(range(nd - 5, nd + 1)))
    imp = ImposedDisplacement([1], lambda t: 1e-4 * t)
    conf = BeamsConfiguration(conf.nodes, conf.beams, None, BoundaryConditions(nd, enc, imp))
    return conf, params, inter


def reference_net(sdf_n=None, drop=0.0):
    """The reference's net (tests/golden: test/input/pos_net.txt, conn_net.txt) as its contact
    tests set it up (test/test_beams2rigid_contact_1_{analytical,discrete}.jl), on the analytical
    sphere or an SDF of it; `drop` lowers it towards the sphere (0.0049: touching)."""
    pos = np.loadtxt(os.path.join(GOLDEN, "pos_net.txt"))
    conn = np.loadtxt(os.path.join(GOLDEN, "conn_net.txt"), dtype=np.int64)
    pos[:, 1] -= drop
    return configs.reference_net(pos, conn, sdf_n=sdf_n)


SETUPS = {
    "test_beams": lambda: configs.reference_test_beams(0.15),
    "contact_sphere": lambda: configs.reference_test_contact_2(0.2),
    "l_cantilever": configs.l_cantilever,
    "ring_sdf": configs.ring_drop,
    "stent_sdf": small_stent,
    "stent_clamped": clamped_stent,
    "reference_net_sphere": lambda: reference_net(drop=0.0049),
    "reference_net_sdf": lambda: reference_net(sdf_n=40, drop=0.0049),
}


def randomize(engines, conf, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    N = len(conf.nodes)
    vals = {}
    for f, s in [("U", 1e-3), ("UDOT", 1e-2), ("UDDOT", 1e-1), ("WDOT", 1e-2), ("WDDOT", 1e-1),
                 ("U_N", 1e-3), ("UDOT_N", 1e-2), ("UDDOT_N", 1e-1), ("WDOT_N", 1e-2), ("WDDOT_N", 1e-1)]:
        vals[f] = rng.standard_normal(3 * N) * s * scale
    for f, s in [("R", 0.05), ("DR", 0.01), ("R_N", 0.05)]:
        R = Rotation.from_rotvec(rng.standard_normal((N, 3)) * s * scale).as_matrix()
        vals[f] = np.transpose(R, (0, 2, 1)).reshape(-1)
    nd = conf.ndofs
    for f in ("FEXT", "FEXT_N", "TINT_N", "TC_N", "TK_N", "D", "DDOT", "DDDOT"):
        vals[f] = rng.standard_normal(nd) * 1e-2
    imposed = rng.standard_normal(len(conf.bcs.fixed_dofs)) * 1e-3
    for e in engines:
        for k, v in vals.items():
            e.set(k, v)
        if e.sizes.n_fixed:
            e.set("IMPOSED", imposed)
    return rng


def pair(name):
    conf, params, inter = SETUPS[name]()
    return conf, params, inter, HipEngine(conf, params, inter), O.OracleEngine(conf, params, inter)


@pytest.mark.parametrize("name", list(SETUPS))
def test_pattern_matches_sparsity_beams(name):
    conf, params, inter, g, o = pair(name)
    pg, po = g.pattern(), o.pattern()
    for a, b in zip(pg, po):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", list(SETUPS))
def test_assembly_parity(name):
    conf, params, inter, g, o = pair(name)
    randomize((g, o), conf, 7)
    g.assemble()
    o.assemble()
    for f in ("TINT", "TK", "TC", "K", "C", "M"):
        assert rel_err(g.get(f), o.get(f)) <= TOL, f
    np.testing.assert_allclose(g.get("ENERGY"), o.get("ENERGY"), rtol=1e-12)
    if inter is not None:
        g.assemble_contact(params.alpha)
        o.assemble_contact(params.alpha)
        for f in ("TC", "K", "C", "M"):
            assert rel_err(g.get(f), o.get(f), floor=1e-300) <= TOL, f
        e_g, e_o = g.get("ENERGY"), o.get("ENERGY")
        assert e_g[2] == pytest.approx(e_o[2], rel=1e-10, abs=1e-300)


@pytest.mark.parametrize("name", list(SETUPS))
def test_tangent_residual_bc_and_extraction_parity(name):
    conf, params, inter, g, o = pair(name)
    randomize((g, o), conf, 11)
    dt, a, b, c = params.initial_timestep, params.alpha, params.beta, params.gamma
    for e in (g, o):
        e.assemble()
        if inter is not None:
            e.assemble_contact(a)
        e.compute_tangent_and_residuals_corrector(dt, a, b, c)
    assert rel_err(g.get("KTAN"), o.get("KTAN")) <= TOL
    assert rel_err(g.get("RES"), o.get("RES")) <= TOL
    for e in (g, o):
        e.apply_boundary_conditions()
        e.extract_free_dofs()
    assert rel_err(g.get("RES"), o.get("RES")) <= TOL
    assert rel_err(g.get("DELTA_D"), o.get("DELTA_D"), floor=1e-300) <= TOL
    assert rel_err(g.get("RES_FREE"), o.get("RES_FREE")) <= TOL
    assert rel_err(g.get("KTAN_FREE"), o.get("KTAN_FREE")) <= TOL
    # predictor tangent/residual from the step-n matrices
    for e in (g, o):
        e.update_converged()
        e.compute_tangent_and_residuals_predictor(dt, a, b, c)
    assert rel_err(g.get("KTAN"), o.get("KTAN")) <= TOL
    assert rel_err(g.get("RES"), o.get("RES")) <= TOL


@pytest.mark.parametrize("name", list(SETUPS))
def test_linear_solve_parity(name):
    conf, params, inter, g, o = pair(name)
    randomize((g, o), conf, 13)
    dt, a, b, c = params.initial_timestep, params.alpha, params.beta, params.gamma
    for e in (g, o):
        e.assemble()
        e.compute_tangent_and_residuals_corrector(dt, a, b, c)
        e.apply_boundary_conditions()
        e.extract_free_dofs()
    its = g.solve_free_dofs()
    o.solve_free_dofs()
    # residual of the device solution measured with the oracle's Ktan_free (CSC, free dofs)
    colptr, rowval, smap, sfree = o.pattern()
    free = np.setdiff1d(np.arange(1, conf.ndofs + 1), conf.bcs.fixed_dofs)
    import scipy.sparse as sp
    K = sp.csc_matrix((o.get("KTAN"), rowval - 1, colptr - 1), shape=(conf.ndofs, conf.ndofs))
    Kf = K[free - 1][:, free - 1]
    x = g.get("DELTA_D_FREE")
    r = o.get("RES_FREE")
    # BiCGStab's relative-residual target (its recursive residual; the true one may differ by rounding)
    assert np.linalg.norm(Kf @ x - r) <= 2 * params.krylov_tol * np.linalg.norm(r)
    assert rel_err(x, o.get("DELTA_D_FREE")) <= 1e-8
    assert rel_err(g.get("DELTA_D"), o.get("DELTA_D")) <= 1e-8
    assert its >= 1
    if name in ("ring_sdf", "stent_sdf", "stent_clamped", "l_cantilever", "test_beams"):
        # open chains, and the ring through its closure correction: the preconditioner is the
        # inverse, so one BiCGStab iteration ends the solve
        assert its == 1, (name, its)


@pytest.mark.parametrize("name", list(SETUPS))
def test_updates_and_norms_parity(name):
    conf, params, inter, g, o = pair(name)
    randomize((g, o), conf, 17)
    dt, a, b, c = params.initial_timestep, params.alpha, params.beta, params.gamma
    rng = np.random.default_rng(3)
    dD = rng.standard_normal(conf.ndofs) * 1e-3
    for e in (g, o):
        e.set("DELTA_D", dD)
        e.compute_global_corrector(dt, b, c)
        e.compute_local_corrector(dt, b, c)
    for f in ("D", "DDOT", "DDDOT", "U", "UDOT", "UDDOT", "WDOT", "WDDOT", "R", "DR"):
        assert rel_err(g.get(f), o.get(f)) <= TOL, f
    free = np.setdiff1d(np.arange(1, conf.ndofs + 1), conf.bcs.fixed_dofs) - 1
    for e in (g, o):
        e.assemble()
        e.compute_tangent_and_residuals_corrector(dt, a, b, c)
        # the reference computes the norms from r_free and ΔD_free, filled by
        # extract_free_dofs! and the solve (utils_solver.jl:94-120)
        e.extract_free_dofs()
        e.set("DELTA_D_FREE", dD[free])
    rg, dg = g.compute_norms_corrector()
    ro, do = o.compute_norms_corrector()
    assert rg == pytest.approx(ro, rel=1e-12) and dg == pytest.approx(do, rel=1e-12)
    for e in (g, o):
        e.initialize_global_predictor()
        e.set("DELTA_D", dD)
        e.compute_global_predictor(dt, b, c)
        e.compute_local_predictor(dt, b, c)
    for f in ("D", "DDOT", "DDDOT", "U", "W", "WDOT", "WDDOT", "R", "DR"):
        assert rel_err(g.get(f), o.get(f)) <= TOL, f
    for e in (g, o):
        e.initialize_global_corrector()
    for f in ("D", "DDOT", "DDDOT", "DELTA_D"):
        assert rel_err(g.get(f), o.get(f), floor=1e-300) <= TOL, f


@pytest.mark.parametrize("name,steps", [("test_beams", 60), ("contact_sphere", 20), ("l_cantilever", 40),
                                        ("ring_sdf", 60), ("stent_sdf", 15), ("stent_clamped", 10),
                                        ("reference_net_sphere", 15), ("reference_net_sdf", 15)])
def test_trajectory_parity(name, steps):
    conf, params, inter = SETUPS[name]()
    g = HipEngine(conf, params, inter)
    o = O.OracleEngine(conf, params, inter)
    sg = run_simulation(conf, params, inter, engine=g, max_steps=steps)
    so = run_simulation(conf, params, inter, engine=o, max_steps=steps)
    assert sg.steps == so.steps == steps
    assert sg.iterations_per_step == so.iterations_per_step
    assert sg.failed_steps == so.failed_steps
    # kinematics to 1e-10; the internal/contact force vectors of a converged step are the small
    # remainder of stiffness × displacement terms that cancel, so the device BiCGStab and the
    # oracle's pivoted LU (both accurate to ε·κ) leave them agreeing to 1e-7 relative
    errs = {f: rel_err(g.get(f), o.get(f), floor=1e-30) for f in ("U_N", "UDOT_N", "R_N", "TINT_N", "TC_N")}
    print(name, steps, errs)
    for f in ("U_N", "R_N"):
        assert errs[f] <= 1e-10, (f, errs[f])
    assert errs["UDOT_N"] <= 1e-9, errs["UDOT_N"]
    for f in ("TINT_N", "TC_N"):
        assert errs[f] <= 1e-7, (f, errs[f])


def test_reference_known_answer_on_gpu():
    """test/test_beams.jl:113-119: tip uⁿ[3] ≈ -0.0100 at t = 0.15 (1500 steps), on the device."""
    conf, params, inter = configs.reference_test_beams(0.15)
    g = HipEngine(conf, params, inter)
    st = run_simulation(conf, params, inter, engine=g)
    assert st.steps == 1500
    assert g.get("U_N").reshape(-1, 3)[-1, 2] == pytest.approx(-0.0100, abs=1e-2)


@pytest.mark.parametrize("builder", ["braided_stent", "net_on_sphere"])
def test_full_size_assembly_parity(builder):
    """Full-size structures (4800 and 3120 beams): one assembly + contact from a perturbed state."""
    conf, params, inter = getattr(configs, builder)()
    g, o = HipEngine(conf, params, inter), O.OracleEngine(conf, params, inter)
    randomize((g, o), conf, 5, scale=0.2)
    for e in (g, o):
        e.assemble()
        e.assemble_contact(params.alpha)
    for f in ("TINT", "TK", "TC", "K", "C", "M"):
        assert rel_err(g.get(f), o.get(f), floor=1e-300) <= TOL, f


IN_CONTACT = {
    # the stent expanded to the vessel wall (radius 2.96 mm in the 3 mm vessel, wire radius 0.05 mm)
    "braided_stent": (lambda: configs.braided_stent(crimp_radius=2.96e-3), 0.9),
    # the 40×40 net pressed 2 cm into the sphere at its centre (a disc of about 0.16 m radius in contact)
    "net_on_sphere": (lambda: configs.net_on_sphere(gap=-0.02), 0.01),
}


@pytest.mark.parametrize("builder", list(IN_CONTACT))
def test_full_size_assembly_parity_in_contact(builder):
    """Full-size structures in contact (4800 and 3120 beams): assembly + contact from the
    configured state with small random perturbations of positions (1e-3 of the wire radius),
    velocities and rotations; the contact path runs on most of the stent's beams and on the
    net's centre."""
    build, min_share = IN_CONTACT[builder]
    conf, params, inter = build()
    g, o = HipEngine(conf, params, inter), O.OracleEngine(conf, params, inter)
    rng = np.random.default_rng(17)
    N = len(conf.nodes)
    rb = float(conf.beams.props[0][0])
    vals = {"U": g.get("U") + rng.standard_normal(3 * N) * 1e-3 * rb,
            "UDOT": rng.standard_normal(3 * N) * 1e-3, "WDOT": rng.standard_normal(3 * N) * 1e-2}
    R0 = g.get("R").reshape(N, 3, 3).transpose(0, 2, 1)
    dR = Rotation.from_rotvec(rng.standard_normal((N, 3)) * 1e-4).as_matrix()
    vals["DR"] = np.transpose(dR, (0, 2, 1)).reshape(-1)
    vals["R"] = np.transpose(dR @ R0, (0, 2, 1)).reshape(-1)
    for e in (g, o):
        for k, v in vals.items():
            e.set(k, v)
        e.assemble()
        e.assemble_contact(params.alpha)
    listed, points = g.contact_stats()
    assert listed >= min_share * len(conf.beams), (listed, len(conf.beams))
    for f in ("TINT", "TK", "TC", "K", "C", "M"):
        assert rel_err(g.get(f), o.get(f), floor=1e-300) <= TOL, f
    assert np.linalg.norm(o.get("TC")) > 0
    e_g, e_o = g.get("ENERGY"), o.get("ENERGY")
    assert e_g[2] == pytest.approx(e_o[2], rel=1e-10)


def test_predictor_reuses_last_factorisation():
    """With Δt unchanged, the predictor's tangent (1+α)Kⁿ + γ/(βΔt)Cⁿ + Mⁿ/(βΔt²) is the last
    corrector iteration's tangent bit for bit, so its chain factors are reused: one
    factorisation fewer than linear solves per such step, and the trajectory still follows the
    oracle step for step."""
    from endobeams_amd import Stepper
    conf, params, inter = SETUPS["l_cantilever"]()
    g = HipEngine(conf, params, inter)
    o = O.OracleEngine(conf, params, inter)
    sg, so = Stepper(g, params), Stepper(o, params)
    sg.initialize()
    so.initialize()
    g.profile_enable(["precond"])
    solves0 = int(g.counters()[2])
    # a predictor can reuse when the attempt before it converged with the same step size
    eligible, last = 0, None
    while sg.stats.steps < 30:
        dt = sg.dt
        assert so.dt == dt
        if last is not None and last == (True, dt):
            eligible += 1
        ok_g, ok_o = sg.advance(), so.advance()
        assert ok_g == ok_o
        last = (ok_g, dt)
    _, factorisations = g.profile_read("precond")
    solves = int(g.counters()[2]) - solves0
    assert sg.stats.iterations_per_step == so.stats.iterations_per_step
    assert sg.stats.failed_steps == so.stats.failed_steps
    assert eligible >= 5, eligible
    assert factorisations == solves - eligible, (factorisations, solves, eligible)
    for f in ("U_N", "R_N"):
        assert rel_err(g.get(f), o.get(f), floor=1e-30) <= 1e-10, f
