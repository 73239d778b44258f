"""The learned warm start on the device: the network's forward pass and gradient against the
numpy restatement (oracle/mlp.py), training on the device, and the solver with the network
against the solver without it (same tolerance, fewer or equal Newton iterations)."""
import numpy as np
import pytest

import mlp as M
from conftest import rel_err

pytestmark = pytest.mark.gpu

SMALL = (27, 32, 16, 6)


def test_forward_matches_numpy():
    from endobeams_amd.learned import WIDTHS, Network
    rng = np.random.default_rng(0)
    for widths in (SMALL, WIDTHS):
        net = Network(widths, seed=1)
        p = rng.normal(0, 0.2, net.n_params).astype(np.float32)
        net.params = p
        X = rng.normal(size=(130, widths[0])).astype(np.float32)
        Y = net.forward(X)
        ref = M.forward(p.astype(np.float64), widths, X.astype(np.float64))[-1]
        assert rel_err(Y, ref) <= 1e-5, rel_err(Y, ref)
        net.close()


def test_gradient_matches_numpy():
    from endobeams_amd.learned import WIDTHS, Network
    rng = np.random.default_rng(2)
    for widths, rows in ((SMALL, 70), (WIDTHS, 300)):
        net = Network(widths, seed=3)
        p = rng.normal(0, 0.2, net.n_params).astype(np.float32)
        net.params = p
        X = rng.normal(size=(rows, widths[0])).astype(np.float32)
        T = rng.normal(size=(rows, widths[-1])).astype(np.float32)
        loss, g = net.loss_gradient(X, T)
        lref, gref = M.loss_and_gradient(p.astype(np.float64), widths, X.astype(np.float64), T.astype(np.float64))
        assert abs(loss - lref) <= 1e-5 * abs(lref)
        assert rel_err(g, gref) <= 1e-4, rel_err(g, gref)
        # every layer's block, not only the norm of the whole vector
        o = 0
        for i, j in zip(widths[:-1], widths[1:]):
            n = i * j + j
            assert rel_err(g[o:o + n], gref[o:o + n]) <= 1e-4
            o += n
        net.close()


def test_training_fits_a_smooth_map():
    from endobeams_amd.learned import Network
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, size=(8192, 27)).astype(np.float32)
    A = rng.normal(size=(27, 6)) / np.sqrt(27)
    T = np.tanh(X @ A).astype(np.float32) + 0.1 * (X[:, :6] ** 2).astype(np.float32)
    net = Network(SMALL, seed=5)
    losses = net.train(X, T, epochs=30, batch=512, lr=3e-3, seed=6)
    print("loss per epoch", losses[[0, 9, 19, 29]])
    assert losses[-1] < 0.1 * losses[0]
    net.close()


def small_stent(seed):
    from endobeams_amd import configs
    return configs.braided_stent(n_wires=6, el_per_wire=32, length=8e-3, sdf_n=64, seed=seed, dt=1.25e-4)


def run(build, ws=None, steps=30, zero=False):
    from endobeams_amd import HipEngine, Stepper
    conf, params, inter = build
    g = HipEngine(conf, params, inter)
    if ws is not None:
        ws.install(g)
    st = Stepper(g, params)
    st.initialize()
    while st.stats.steps < steps:
        st.advance()
    return g, st.stats


def test_zero_correction_reproduces_the_reference_predictor():
    """A network whose output is zero makes the learned start the reference predictor's start
    (ties keep the network's): the run, with its snapshot/restore and per-instance choice, must be
    the run without a network, bit for bit."""
    from endobeams_amd import learned as L
    net = L.Network(L.WIDTHS, seed=0)
    net.params = np.zeros(net.n_params, np.float32)
    ws = L.WarmStart(net, np.ones(27, np.float32), np.ones(6, np.float32))
    g0, s0 = run(small_stent(31), steps=12)
    g1, s1 = run(small_stent(31), ws, steps=12)
    assert s0.iterations_per_step == s1.iterations_per_step and s0.failed_steps == s1.failed_steps
    for f in ("U_N", "R_N", "UDOT_N", "TINT_N", "TC_N"):
        assert np.array_equal(g0.get(f), g1.get(f)), f
    stats = L.predictor_stats(g1)
    assert stats[1] == 0 and stats[0] == 12 + s1.failed_steps - 3


def test_learned_warm_start_against_the_reference_predictor():
    """Trained on three seeded geometries, run on a fourth: the solver with the network converges
    every step to the same tolerance. On these structures the network's start does not cut Newton
    iterations (the corrector converges linearly with the reference's tangent, and only a nearly
    exact start would save iterations; the measured outcome is printed). What is pinned: the run
    with the network takes no more Newton iterations than the run without it, and when the
    network's start loses every evaluation, each step is the reference predictor's, bit for bit."""
    from endobeams_amd import HipBatchEngine
    from endobeams_amd import learned as L
    train = [small_stent(s) for s in (11, 12, 13)]
    eng = HipBatchEngine(train)
    X, T = L.collect(eng, train[0][1], 30)
    eng.close()
    ws = L.WarmStart.train(X, T, epochs=40, batch=1024, lr=1e-3, seed=0)
    print("samples", X.shape, "loss", ws.losses[[0, -1]])
    assert ws.losses[-1] < 0.5 * ws.losses[0]
    gp, sp = run(small_stent(21))
    gl, sl = run(small_stent(21), ws)
    stats = L.predictor_stats(gl)
    ip, il = np.mean(sp.iterations_per_step[3:]), np.mean(sl.iterations_per_step[3:])
    print(f"newton/step: reference predictor {ip:.2f} ({sp.failed_steps} halvings), with the network {il:.2f} "
          f"({sl.failed_steps} halvings); network start kept {stats[0]} of {stats.sum()}")
    assert sl.steps == sp.steps == 30
    assert stats.sum() == 30 + sl.failed_steps - 3
    # the network may only keep a start that costs no iteration over the run
    assert il <= ip, (il, ip)
    if stats[0] == 0:
        assert sl.iterations_per_step == sp.iterations_per_step and sl.failed_steps == sp.failed_steps
        for f in ("U_N", "R_N", "TC_N"):
            np.testing.assert_array_equal(gl.get(f), gp.get(f), err_msg=f)


def test_solver_inference_path_matches_numpy():
    """The network as the solver runs it (inputs formed on the device, inference into the
    context's own layer buffers on the solver's stream, the correction added on free dofs) against
    the numpy forward pass of the same random non-zero parameters on the same inputs, to 1e-5."""
    from endobeams_amd import HipEngine, Stepper
    from endobeams_amd import learned as L
    build = small_stent(41)
    # input scales from the inputs of this trajectory, so that every layer works on O(1) values
    b = HipEngine(*build)
    sb = Stepper(b, build[1])
    sb.initialize()
    while sb.stats.steps < 4:
        sb.advance()
    Xs = L.sample(b)[0].astype(np.float64)
    in_scale = (1.0 / np.maximum(Xs.std(axis=0), 1e-30)).astype(np.float32)
    out_scale = (0.1 * np.maximum(Xs[:, :6].std(axis=0), 1e-30)).astype(np.float32)
    rng = np.random.default_rng(8)
    net = L.Network(L.WIDTHS, seed=9)
    params = rng.normal(0, 0.3, net.n_params).astype(np.float32)
    net.params = params
    ws = L.WarmStart(net, in_scale, out_scale)
    engines = []
    for with_net in (False, True):
        e = HipEngine(*build)
        if with_net:
            ws.install(e)
            L.set_backoff(e, 0)
        st = Stepper(e, build[1])
        st.initialize()
        while st.stats.steps < 3:
            st.advance()
        e.predictor(st.dt)
        engines.append(e)
    plain, learned = engines
    p = plain.get("DELTA_D")                  # the reference predictor's increment of this step
    inc = L.learned_increment(learned)        # the network's start: p + out_scale ⊙ f(in_scale ⊙ x)
    X = L.sample(learned)[0].astype(np.float64)
    y = M.forward(params.astype(np.float64), L.WIDTHS, X * in_scale.astype(np.float64))[-1]
    want = y * out_scale.astype(np.float64)
    got = (inc - p).reshape(-1, 6)
    assert np.linalg.norm(want) > 0
    assert rel_err(got, want) <= 1e-5, rel_err(got, want)
