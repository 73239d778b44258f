"""Newton iterations replayed from a captured HIP graph (eb_set_exec_mode) and launch shapes chosen
by the autotuner (eb_autotune) leave every result unchanged: the same trajectories, bit for bit,
with the same Newton iteration counts, as the eagerly launched path (which the parity suite checks
against the oracle); the one exception is the chain preconditioner's algorithm, which changes the
linear solve's rounding only. The graph is replayed many times per step with the step's eager work (the
predictor, commits, rollbacks on Δt halvings) and host synchronisations between the replays."""
import numpy as np
import pytest

from conftest import rel_err
from endobeams_amd import HipBatchEngine, HipEngine, Stepper, abi, configs

pytestmark = pytest.mark.gpu

# the two chain algorithms round differently inside the linear solve, which ends at a relative
# residual of 1e-11: states after a few steps agree far below the Newton tolerance (1e-8)
ALGO_TOL = 1e-9

FIELDS = ("U", "UDOT", "UDDOT", "W", "WDOT", "WDDOT", "R", "DR", "TINT", "TC", "K_N", "ENERGY")


def run(engine, params, steps, setup=None):
    st = Stepper(engine, params)
    st.initialize()
    if setup is not None:
        setup(engine, st)
    ks, attempts = [], 0
    while st.stats.steps < steps and attempts < 4 * steps:
        attempts += 1
        st.advance()
    ks = list(st.stats.iterations_per_step)
    return ks, {f: engine.get(f).copy() for f in FIELDS}, st


def stents():
    shapes = [(4, 24, 0), (6, 32, 1), (8, 28, 2)]
    return [configs.braided_stent(n_wires=w, el_per_wire=m, length=8e-3, sdf_n=48, seed=s) for (w, m, s) in shapes]


CASES = {
    "ring_sdf_fp64": lambda: (configs.ring_drop(), abi.EB_PRECISION_FP64, 12),
    "stent_fp64": lambda: (configs.braided_stent(n_wires=4, el_per_wire=40, sdf_n=64), abi.EB_PRECISION_FP64, 8),
    "l_cantilever_fp32": lambda: (configs.l_cantilever(), abi.EB_PRECISION_FP32, 12),
}


def make(case):
    (conf, params, inter), prec, steps = CASES[case]()
    return HipEngine(conf, params, inter, precision=prec), params, steps


@pytest.mark.parametrize("case", list(CASES))
def test_graph_replays_equal_eager_launches(case):
    eager, params, steps = make(case)
    k0, f0, _ = run(eager, params, steps)
    eager.close()
    g, params, steps = make(case)
    g.set_exec_mode(True)
    k1, f1, _ = run(g, params, steps)
    mode, captures, replays = g.exec_stats()
    print(f"{case}: iterations {k1}, graph captures {captures}, replays {replays}", flush=True)
    assert mode == abi.EB_EXEC_GRAPH and captures >= 1 and replays >= sum(k1)
    assert k1 == k0
    for f in FIELDS:
        assert np.array_equal(f0[f], f1[f]), (f, np.max(np.abs(f0[f] - f1[f])))
    g.close()


@pytest.mark.parametrize("precision", [abi.EB_PRECISION_FP64, abi.EB_PRECISION_FP32])
def test_graph_batch_equal_eager(precision):
    """A batch (converged instances frozen on the device while the others iterate) in graph mode."""
    eager = HipBatchEngine(stents(), precision=precision)
    k0, f0, _ = run(eager, eager.params, 6)
    eager.close()
    g = HipBatchEngine(stents(), precision=precision)
    g.set_exec_mode(True)
    k1, f1, _ = run(g, g.params, 6)
    print("batch iterations", k1, "exec stats", g.exec_stats(), flush=True)
    assert k1 == k0
    for f in FIELDS:
        assert np.array_equal(f0[f], f1[f]), f
    g.close()


def test_graph_timing_events():
    """Kernel timings in graph mode come from event nodes of the graph."""
    eng, params, steps = make("stent_fp64")
    eng.set_exec_mode(True)

    def enable(e, st):
        e.profile_enable(["beam", "precond"])
    k, _, _ = run(eng, params, 4, setup=enable)
    for kern in ("beam", "precond"):
        ms, n = eng.profile_read(kern)
        print(kern, ms, n, flush=True)
        assert n >= sum(k) and ms > 0
    eng.close()


def test_autotune_choices_and_unchanged_results():
    (conf, params, inter), prec, steps = CASES["stent_fp64"]()
    eager = HipEngine(conf, params, inter, precision=prec)
    k0, f0, _ = run(eager, params, 6)
    eager.close()
    tuned = HipEngine(conf, params, inter, precision=prec)
    with pytest.raises(Exception):
        tuned.autotune()          # no tangent yet
    picked = {}

    def tune_after_first_step(e, st):
        st.advance()
        choice, times = e.autotune()
        picked.update(choice)
        print("autotune", choice, times, flush=True)
        for i, name in enumerate(abi.TUNE_NAMES):
            assert choice[name] in abi.TUNE_CANDIDATES[i]
            assert times[name] and all(v > 0 for v in times[name].values())
            assert choice[name] == min(times[name], key=times[name].get)
    k1, f1, _ = run(tuned, params, 6, setup=tune_after_first_step)
    assert tuned.tuning == picked
    assert k1 == k0
    for f in FIELDS:
        if picked["chain_algo"] == abi.EB_CHAIN_BLOCK_LU:   # the default algorithm: bit-identical
            assert np.array_equal(f0[f], f1[f]), f
        else:
            assert rel_err(f1[f], f0[f], floor=1e-30) <= ALGO_TOL, f
    # every launch shape, forced, gives the same trajectory bit for bit (graph mode); the shapes
    # of the cyclic-reduction kernels are run with that algorithm
    refs = {}
    for algo in abi.TUNE_CANDIDATES[abi.TUNE_NAMES.index("chain_algo")]:
        ref = HipEngine(conf, params, inter, precision=prec)
        ref.tuning = {"chain_algo": algo}
        refs[algo] = run(ref, params, 3)
        ref.close()
    for i, name in enumerate(abi.TUNE_NAMES):
        for c in abi.TUNE_CANDIDATES[i]:
            algo = c if name == "chain_algo" else abi.EB_CHAIN_CYCLIC_REDUCTION
            e = HipEngine(conf, params, inter, precision=prec)
            e.tuning = {name: c, "chain_algo": algo}
            e.set_exec_mode(True)
            kc, fc, _ = run(e, params, 3)
            k3, f3, _ = refs[algo]
            assert kc == k3, (name, c)
            for f in FIELDS:
                assert np.array_equal(f3[f], fc[f]), (name, c, f)
            e.close()
    # the two chain algorithms: the same iteration counts, states within the solve's rounding
    (ka, fa, _), (kb, fb, _) = refs[abi.EB_CHAIN_CYCLIC_REDUCTION], refs[abi.EB_CHAIN_BLOCK_LU]
    assert ka == kb
    errs = {f: rel_err(fb[f], fa[f], floor=1e-30) for f in FIELDS}
    print("block LU vs cyclic reduction", errs, flush=True)
    assert max(errs.values()) <= ALGO_TOL, errs
    tuned.close()
