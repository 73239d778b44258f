#!/usr/bin/env python3
"""Benchmark: element-steps/s and Newton iterations/step of the device Newton iteration.

One "step" is one HHT-α time step (predictor + Newton corrector to residual and displacement
tolerance 1e-8, the reference's adaptive Δt), inputs resident on the GPU. Workload: the braided
stent, 24 helical wires × 200 beams (4800 beams, 4824 nodes, 28 944 dofs), crimped to 1.5 mm and
released in a rigid vessel given as a 256³ FP32 signed distance field.

Four forms are timed in the same run:
  plain       — FP64 element/contact kernels, one stent per context, the reference predictor;
  fp64_batch  — the same FP64 kernels on a batch of seeded stent variants advanced in one context
                (each instance its own vessel SDF and its own time step);
  fast        — the precision form (FP32 element matrices and contact tangent; forces, residuals,
                rotations, Krylov and updates FP64) on the same batch;
  learned     — the fast form with the learned warm start: a 3×128 FP32 network trained in this
                run, by the library's own kernels, on trajectories of 8 other seeded stents; the
                reference predictor stays the per-instance fallback.
The batch size is the one the batch-shape autotuner picks (`--batch-candidates`, contexts of each
size timed over their first steps). `value` is the fast form's whole-job throughput: the learned
start is timed beside it and not quoted, because on these stents its start saves no Newton
iteration over the reference predictor's (it costs its evaluations). The other forms are
reported with the ratios (the FP64 batch separates batching from the FP32 element matrices). Also reported: the roofline of the element kernel (algorithmic flops per launch ÷ average
launch time from HIP events in this run), HBM roofline records of the matrix gather and the
contact kernel (algorithmic bytes per launch ÷ average launch time), the CPU restatement's
throughput on a bounded sample of the plain workload (one thread and all OpenMP threads) and the
fast form's trajectory error against that FP64 CPU run at the same step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "endobeams.jl_amd"))

METRIC = "element-steps/s & Newton iters/step, learned+FP32 vs plain FP64 GPU, same tol"

# FP64 operations (+, -, ×, ÷; transcendentals and sqrt excluded) of one compute_beams call with
# 3 Gauss points and no damping, counted by the op-counting build of the CPU restatement
# (tests/test_flop_count.py keeps this figure in sync).
FLOPS_PER_BEAM = 26547
# The same count without the operations on exact zeros (x·0, x ± 0, 0 / x): the structural zeros of
# the dense 3x3 algebra, which the element kernel's build folds away. More than half of the dense
# count multiplies or adds such zeros; this is the work the kernel still does.
FLOPS_PER_BEAM_NONZERO = 12451
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector
PEAK_HBM_GBS = 8000.0


def roofline(kernel, avg_s, beams, peak, traffic=None, peak_fp64=PEAK_FP64_TFLOPS):
    """The element kernel's roofline record: the dense algorithmic flops per launch (beams: those of
    the instances active in the launch, averaged over the timed launches; and, beside them, the
    flops left once structural zeros are folded) over the measured average launch. The
    kernel runs on the vector ALUs (no matrix cores), part of it in FP64 in the precision form:
    the fraction is given against the FP32 vector peak and, beside it, against the FP64 one."""
    achieved = FLOPS_PER_BEAM * beams / avg_s / 1e12
    nonzero = FLOPS_PER_BEAM_NONZERO * beams / avg_s / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "frac_of_fp64_peak": achieved / peak_fp64, "traffic": traffic, "kernel": kernel,
            "avg_launch_us": avg_s * 1e6, "beams_per_launch": beams, "flops_per_launch": FLOPS_PER_BEAM * beams,
            "nonzero_flops_per_launch": FLOPS_PER_BEAM_NONZERO * beams, "nonzero_achieved": nonzero,
            "nonzero_frac": nonzero / peak}


# Algorithmic HBM bytes of the HBM-bound kernels (s: bytes of the staging type, 4 in the precision
# form, 8 in FP64).
def element_kernels(reduced, contact, beams):
    """The element kernel's launches for one assembly of `beams` beams: the precision form runs two
    launches of two parts each from 2 048 tiles of 64 beams on, one launch of all four below."""
    if reduced and -(-beams // 64) >= 2048:
        return ("beam_detect_kernel_half<float>" if contact else "beam_kernel_half<float, 0>") + \
            " + beam_kernel_half<float, 2>"
    t = "float" if reduced else "double"
    return f"beam_detect_kernel<{t}>" if contact else f"beam_kernel<{t}>"


def gather_bytes(E, nnz, s):
    """gather_matrices forming the tangent: each beam's three 144-entry staging records read once
    (plus its four block-source entries), every tangent entry written once in FP64, and per 6×6
    block its pointer and instance entries."""
    return E * (3 * 144 * s + 4 * 4) + nnz * 8 + (nnz / 36) * 8


def contact_beam_bytes(s):
    """contact_kernel per listed beam: two nodes' X₀, u, u̇, ẇ, R, ΔR (30 doubles each), the beam's
    l₀, Rₑ⁰, radius and node indices, its list entry, read-modify-write of its K and C staging
    records (2 × 144 entries), Tᶜ (12 doubles) and the contact energy."""
    return 2 * 30 * 8 + (11 * 8 + 8) + 8 + 2 * 144 * s * 2 + 12 * 8 + 8


def hbm_record(kernel, bytes_per_launch, avg_s, **extra):
    achieved = bytes_per_launch / avg_s / 1e9
    return {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_s * 1e6,
            **extra}


def cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def stent(seed=None, sdf_n=256):
    from endobeams_amd import configs
    return configs.braided_stent(sdf_n=sdf_n, seed=seed)


def run_steps(stepper, n):
    """Advances until n more steps have converged; returns (converged, attempts)."""
    done = attempts = 0
    while done < n and not stepper.finished:
        attempts += 1
        if stepper.advance():
            done += 1
    return done, attempts


PROFILED = ("beam", "contact", "gather", "gather_matrices", "precond", "krylov", "update")


def _kernel_times(eng, contact0):
    """Per kernel group (total ms, launches) over the timed steps, the contact kernel's work (beams
    listed, Gauss points in contact) and the work the library counted for the timed launches of
    the element kernel, the matrix gather and the chain factorisation (beams, tangent entries,
    chain positions of the instances active in each launch: converged instances are skipped)."""
    out = {k: eng.profile_read(k) for k in PROFILED}
    c1 = eng.contact_stats()
    return out, (c1[0] - contact0[0], c1[1] - contact0[1]), eng.work_counters()


def timed(eng, params, E, warmup, steps, graph=True, autotune=True, profile=True):
    """One structure: warm-up steps (then the autotuner, on the live state), then `steps` timed
    converged steps of the library's own time loop (eb_run_steps). profile: the per-kernel HIP
    events on (they add two stream markers per kernel group to the timed steps)."""
    from endobeams_amd import Stepper, abi
    eng.set_exec_mode(graph)
    st = Stepper(eng, params)
    st.initialize()
    run_steps(st, warmup)
    tuning = eng.autotune() if autotune else (eng.tuning, None)
    if profile:
        eng.profile_enable(list(PROFILED))
    c0 = eng.contact_stats()
    state = abi.RunState(t=st.tn, dt=st.dt, successive=st.successive)
    t0 = time.perf_counter()
    state = eng.run_steps(steps, state)
    wall = time.perf_counter() - t0
    done = state.steps
    kt, contact, work = _kernel_times(eng, c0)
    its = state.newton_iterations / max(done, 1)
    return {"steps": done, "attempts": done + state.failed_steps, "wall": wall, "E": E, "nnz": eng.sizes.nnz,
            "value": E * done / wall, "ms_per_step": 1e3 * wall / max(done, 1),
            "newton_iters_per_step": its, "newton_iters_per_step_max": its,
            "failed_steps": state.failed_steps,
            "wall_s_per_simulated_s": wall / (state.t - st.tn) if state.t > st.tn else None,
            "kernels": kt, "contact": contact, "work": work,
            "exec": {"mode": graph if graph == "fused" else ("graph" if graph else "eager"),
                     "graph_captures_replays": list(eng.exec_stats()[1:]),
                     "tuning": tuning[0], "tuning_ms": tuning[1]}}


def timed_batch(eng, warmup, steps, graph=True, autotune=True):
    """A batch context: every instance with its own Δt (eb_run_steps_batch): `warmup` converged
    steps each (then the autotuner), then `steps` timed converged steps each."""
    eng.set_exec_mode(graph)
    eng.initialize_state_simulation()
    states = eng.run_steps_batch(warmup)
    tuning = eng.autotune() if autotune else (eng.tuning, None)
    eng.profile_enable(list(PROFILED))
    E_s = [int(eng.beam_off[s + 1] - eng.beam_off[s]) for s in range(eng.S)]
    n0 = [r.steps for r in states]
    i0 = [r.newton_iterations for r in states]
    f0 = [r.failed_steps for r in states]
    t_0 = [r.t for r in states]
    c0 = eng.contact_stats()
    t0 = time.perf_counter()
    states = eng.run_steps_batch(steps, states)
    wall = time.perf_counter() - t0
    dn = [r.steps - n for r, n in zip(states, n0)]
    di = [r.newton_iterations - i for r, i in zip(states, i0)]
    sim = [r.t - t for r, t in zip(states, t_0)]
    kt, contact, work = _kernel_times(eng, c0)
    return {"steps": min(dn), "attempts": sum(dn) + sum(r.failed_steps - f for r, f in zip(states, f0)),
            "wall": wall, "E": sum(E_s), "nnz": eng.sizes.nnz,
            "value": sum(e * d for e, d in zip(E_s, dn)) / wall, "ms_per_step": 1e3 * wall / max(min(dn), 1),
            "newton_iters_per_step": sum(di) / max(sum(dn), 1),
            "newton_iters_per_step_max": max(i / max(d, 1) for i, d in zip(di, dn)),
            "failed_steps": sum(r.failed_steps - f for r, f in zip(states, f0)),
            "wall_s_per_simulated_s": wall / (sum(sim) / len(sim)) if sum(sim) > 0 else None,
            "kernels": kt, "contact": contact, "work": work,
            "exec": {"mode": graph if graph == "fused" else ("graph" if graph else "eager"),
                     "graph_captures_replays": list(eng.exec_stats()[1:]),
                     "tuning": tuning[0], "tuning_ms": tuning[1]}}


def kernel_avg_s(form, k):
    ms, n = form["kernels"][k]
    return ms * 1e-3 / n if n else None


def active_per_launch(form, group, which):
    """Units (beams, tangent entries, chain positions) one timed launch of a kernel group worked
    on, averaged over the launches: the library counts those of the instances active in each."""
    n = form["kernels"][group][1]
    return form["work"][which] / n if n else 0.0


def hbm_records(form, s):
    """The HBM roofline records of the matrix gather and the contact kernel of a timed form."""
    out = []
    g = kernel_avg_s(form, "gather_matrices")
    if g:
        entries = active_per_launch(form, "gather_matrices", 1)
        beams = entries * form["E"] / form["nnz"]   # beams of the active instances (same beams per entry)
        out.append(hbm_record("gather_matrices", gather_bytes(beams, entries, s), g,
                              launches=form["kernels"]["gather_matrices"][1], entries_per_launch=entries,
                              entries_all_instances=form["nnz"]))
    c = kernel_avg_s(form, "contact")
    n = form["kernels"]["contact"][1]
    if c and n and form["contact"][0] > 0:
        listed, gps = form["contact"][0] / n, form["contact"][1] / n
        out.append(hbm_record("contact_kernel", listed * contact_beam_bytes(s) + gps * 27 * s, c, launches=n,
                              beams_in_contact_per_launch=listed, gauss_points_in_contact_per_launch=gps))
    return out


def breakdown(form):
    """Device time per time step of each kernel group (HIP events around every launch; the matrix
    gather is part of `gather`), and its share of the wall time."""
    steps = max(form["steps"], 1)
    out = {k: form["kernels"][k][0] / steps for k in PROFILED}
    out["wall"] = form["ms_per_step"]
    return out


def summary(form):
    keys = ("value", "ms_per_step", "newton_iters_per_step", "newton_iters_per_step_max", "failed_steps",
            "wall_s_per_simulated_s")
    out = {k: form[k] for k in keys}
    b = kernel_avg_s(form, "beam")
    out["beam_kernel_us"] = b * 1e6 if b else None
    out["ms_per_step_by_kernel"] = breakdown(form)
    return out


def _cpu_leg(seconds, timing):
    """The FP64 CPU restatement on the plain workload: its first step, then steps for `seconds`."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from endobeams_amd import Stepper
    conf, params, inter = stent()
    eng = O.OracleEngine(conf, params, inter, timing=timing)
    st = Stepper(eng, params)
    st.initialize()
    run_steps(st, 1)
    t0 = time.perf_counter()
    n0, it0 = st.stats.steps, st.stats.newton_iterations
    while time.perf_counter() - t0 < seconds and not st.finished:
        run_steps(st, 1)
    wall = time.perf_counter() - t0
    steps = st.stats.steps - n0
    return len(conf.beams) * steps / wall, steps, st.stats.newton_iterations - it0, wall, eng, st.stats.steps


def cpu_baseline(seconds):
    """The C++ FP64 restatement (a port, not Julia, which is not available): one thread (the
    parity build), then all the cores OpenMP is given (the timing build: element and contact loops
    over threads; the scatter in beam order and the banded-LU direct solve stay serial)."""
    v1, n1, i1, w1, eng, n_done = _cpu_leg(seconds * 0.6, False)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    vn, nn, i_n, wn, _, _ = _cpu_leg(seconds * 0.4, True)
    info = cpu_info()
    out = {"value": vn, "unit": "element-steps/s", "cores": threads, "kind": "port",
           "sample": f"braided stent (4800 beams), {nn} converged steps after its first step, {i_n} Newton "
                     f"iterations, {wn:.1f} s, C++ FP64 restatement with banded-LU direct solve, element and "
                     f"contact loops on {threads} OpenMP threads",
           **info,
           "one_thread": {"value": v1, "cores": 1,
                          "sample": f"{n1} converged steps after the first, {i1} Newton iterations, {w1:.1f} s"}}
    return out, eng, n_done


def sweep(args):
    """The design sweep: `--instances` braided stents of uneven sizes (1 200 to 14 400 beams each),
    each in its own vessel SDF, FP32 element matrices, advanced in batch contexts whose size the
    batch-shape autotuner picks (each context then autotunes its launch shapes); the learned warm
    start (trained on 8 seeded stents) with the reference predictor as the fallback."""
    from endobeams_amd import HipBatchEngine, Stepper, abi, autotune, configs
    from endobeams_amd import learned as L
    t = time.perf_counter()
    inst = configs.stent_sweep(args.instances, sdf_n=args.sweep_sdf_n)
    E_all = sum(len(c.beams) for (c, _, _) in inst)
    print(f"# sweep of {len(inst)} stents, {E_all} beams, built in {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    t = time.perf_counter()
    tr = [stent(seed=1000 + s, sdf_n=args.sweep_sdf_n) for s in range(8)]
    teng = HipBatchEngine(tr, precision=abi.EB_PRECISION_FP32)
    X, T = L.collect(teng, tr[0][1], args.train_steps)
    teng.close()
    ws = L.WarmStart.train(X, T, epochs=args.epochs, batch=4096, lr=1e-3, seed=0)
    print(f"# network trained in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    if args.sub_batch > 0:
        size, rates, shapes = min(args.sub_batch, len(inst)), {}, {}
    else:
        t = time.perf_counter()
        size, rates, shapes = autotune.pick_batch_size(
            inst, [c for c in (64, 128, 256, 512, 1024)], warmup=1, steps=2,
            log=lambda c, r: print(f"#   {c} instances per context: {r:.0f} element-steps/s", file=sys.stderr, flush=True))
        print(f"# batch shape: {size} instances per context ({ {k: round(v) for k, v in rates.items()} }), "
              f"tuned in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    groups = autotune.split(inst, size)
    engs, states = [], []
    for gi, g in enumerate(groups):
        print(f"# context {gi + 1}/{len(groups)}: {len(g)} instances", file=sys.stderr, flush=True)
        e = HipBatchEngine(g, precision=abi.EB_PRECISION_FP32)
        e.set_exec_mode(args.exec == "graph")
        ws.install(e)
        e.initialize_state_simulation()
        states.append(e.run_steps_batch(args.warmup))
        e.autotune()
        e.profile_enable(["beam"])
        engs.append(e)
    base = [[(r.steps, r.newton_iterations, r.failed_steps, r.t) for r in st] for st in states]
    t0 = time.perf_counter()
    for _ in range(args.steps):   # the contexts in turn, one step of each instance (eb_run_steps_batch)
        for i, e in enumerate(engs):
            states[i] = e.run_steps_batch(1, states[i])
    wall = time.perf_counter() - t0
    value = steps_sum = its_sum = failed = 0
    sims = []
    for e, st, b0 in zip(engs, states, base):
        for sidx, (r, (n0, i0, f0, t_0)) in enumerate(zip(st, b0)):
            E_s = int(e.beam_off[sidx + 1] - e.beam_off[sidx])
            value += E_s * (r.steps - n0)
            steps_sum += r.steps - n0
            its_sum += r.newton_iterations - i0
            failed += r.failed_steps - f0
            sims.append(r.t - t_0)
    value /= wall
    its = its_sum / max(steps_sum, 1)
    sim = sum(sims) / len(sims)
    kms = kn = kbeams = 0.0
    for e in engs:
        ms, n = e.profile_read("beam")
        kms += ms
        kn += n
        kbeams += e.work_counters()[0]
    chosen = [L.predictor_stats(e).tolist() for e in engs]
    avg_s = (kms / max(kn, 1)) * 1e-3
    E_launch = kbeams / max(kn, 1)   # beams of the active instances per launch
    out = {
        "metric": METRIC, "value": value, "unit": "element-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"stent design sweep: {len(inst)} braided stents, {E_all} beams, "
                               f"{args.sweep_sdf_n}^3 FP32 vessel SDF each, Newton tolerance 1e-8",
                   "form": "learned warm start + FP32 element matrices + FP64 forces/solver, "
                           f"{len(groups)} batch contexts of {size} instances",
                   "exec": args.exec, "dt": 1e-3, "tolerance": 1e-8, "beams": E_all},
        "newton_iters_per_step": its, "failed_steps": failed,
        "wall_s_per_simulated_s": wall / sim if sim > 0 else None,
        "time_steps": "each stent its own Δt (eb_run_steps_batch)",
        "batch_autotune": {"instances_per_context": size, "element_steps_per_s": rates,
                           "launch_shapes": shapes.get(size)},
        "learned_choices_per_context": chosen,
        "roofline": {**roofline(element_kernels(True, True, sum(len(c.beams) for (c, _, _) in groups[0])), avg_s,
                                E_launch, PEAK_FP32_TFLOPS), "launches": int(kn)},
    }
    for e in engs:
        e.close()
    print(json.dumps(out))


def small_config(args):
    """One of the smaller structures: plain FP64 and the precision form, one structure per
    context, each timed over `--steps` converged steps; the precision form's error against the
    FP64 CPU restatement after `--compare-steps` steps (nodal positions, rotations and contact
    forces); the restatement's own throughput on those steps."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from endobeams_amd import HipEngine, Stepper, abi, configs
    build = {"cantilever": configs.l_cantilever, "ring": configs.ring_drop, "net": configs.net_on_sphere}[args.workload]
    forms = {}
    # one structure whose chain preconditioner holds every coupling: also each step as one launch
    fused = args.workload in ("cantilever", "ring") and args.exec != "graph"
    for name, prec in (("fp64", abi.EB_PRECISION_FP64), ("fp32", abi.EB_PRECISION_FP32)):
        conf, params, inter = build()
        eng = HipEngine(conf, params, inter, precision=prec)
        forms[name] = timed(eng, params, len(conf.beams), args.warmup, args.steps, graph=args.exec == "graph",
                            autotune=not args.no_autotune)
        eng.close()
        print(f"# {args.workload} {name}: {forms[name]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
        if fused:
            conf, params, inter = build()
            eng = HipEngine(conf, params, inter, precision=prec)
            # timed without the per-kernel events: this leg's line is its wall time (the kernel
            # figures come from the eager leg above)
            forms[name + "_fused"] = timed(eng, params, len(conf.beams), args.warmup, args.steps, graph="fused",
                                           autotune=False, profile=False)
            eng.close()
            print(f"# {args.workload} {name}, one launch per step: {forms[name + '_fused']['ms_per_step']:.3f} ms/step",
                  file=sys.stderr, flush=True)
    # the CPU restatement over the GPU legs' timed window: the same structure and start, `warmup`
    # converged steps untimed, then `steps` converged steps timed (failed attempts included, as there)
    conf, params, inter = build()
    ref = O.OracleEngine(conf, params, inter)
    st = Stepper(ref, params)
    st.initialize()
    run_steps(st, args.warmup)
    t0 = time.perf_counter()
    cpu_done, cpu_attempts = run_steps(st, args.steps)
    cpu_wall = time.perf_counter() - t0
    # the comparison starts the ring and the net at the surface, so that contact forces are compared
    near = {"cantilever": configs.l_cantilever, "ring": lambda: configs.ring_drop(height=0.06),
            "net": lambda: configs.net_on_sphere(gap=0.005)}[args.workload]
    conf, params, inter = near()
    ref = O.OracleEngine(conf, params, inter)
    st = Stepper(ref, params)
    st.initialize()
    run_steps(st, args.compare_steps)
    g = HipEngine(conf, params, inter, precision=abi.EB_PRECISION_FP32)
    sg = Stepper(g, params)
    sg.initialize()
    run_steps(sg, args.compare_steps)
    err = {}
    for f in ("U_N", "R_N", "TC_N"):
        a, b = g.get(f), ref.get(f)
        nb = float(np.linalg.norm(b))
        err[f] = float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))
    E = len(conf.beams)
    f32, f64 = forms["fp32"], forms["fp64"]
    avg_s = kernel_avg_s(f32, "beam")
    # the line's value: the faster execution of the precision form (one launch per step where it applies)
    head = forms["fp32_fused"] if fused and forms["fp32_fused"]["value"] > f32["value"] else f32
    out = {
        "metric": METRIC, "value": head["value"], "unit": "element-steps/s", "n_gpus": 1, "steps": head["steps"],
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": {"cantilever": "L-shaped cantilever, 20 beams",
                                "ring": "ring of 60 beams dropped on a plane (64x64x32 FP32 SDF)",
                                "net": "40x40 net of 3120 beams dropped on a sphere (128^3 FP32 SDF)"}[args.workload],
                   "form": "FP32 element matrices + FP64 forces/solver, one structure",
                   "exec": "fused" if head is not f32 else args.exec,
                   "dt": params.initial_timestep, "tolerance": params.tolerance_residual, "beams": E},
        "newton_iters_per_step": f32["newton_iters_per_step"], "failed_steps": f32["failed_steps"],
        "wall_s_per_simulated_s": f32["wall_s_per_simulated_s"],
        "plain_fp64": {"value": f64["value"], "ms_per_step": f64["ms_per_step"],
                       "newton_iters_per_step": f64["newton_iters_per_step"],
                       "wall_s_per_simulated_s": f64["wall_s_per_simulated_s"]},
        "by_execution": {k: {"value": v["value"], "ms_per_step": v["ms_per_step"],
                             "newton_iters_per_step": v["newton_iters_per_step"], "failed_steps": v["failed_steps"]}
                         for k, v in forms.items()},
        "fp32_error_vs_fp64_oracle": {"steps": args.compare_steps,
                                      "start": "at the surface" if args.workload != "cantilever" else "as configured",
                                      **err},
        "cpu_baseline": {"value": E * cpu_done / cpu_wall, "unit": "element-steps/s", "cores": 1, "kind": "port",
                         "ms_per_step": 1e3 * cpu_wall / max(cpu_done, 1),
                         "sample": f"the timed window of the GPU legs: {cpu_done} converged steps after {args.warmup} "
                                   f"(attempts {cpu_attempts}), C++ FP64 restatement, 1 thread"},
        "roofline": {**roofline(element_kernels(True, inter is not None, E), avg_s,
                                active_per_launch(f32, "beam", 0), PEAK_FP32_TFLOPS),
                     "launches": f32["kernels"]["beam"][1]},
        "hbm_rooflines": hbm_records(f32, 4),
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=0,
                    help="stents advanced together by the batched forms (0: picked by the batch-shape autotuner)")
    ap.add_argument("--batch-candidates", default="128,256,512,1024", help="sizes the batch-shape autotuner times")
    ap.add_argument("--sdf-n", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default=None,
                    help="comma-separated forms to time: plain, fp64_batch, fast, learned (default: all)")
    ap.add_argument("--train-steps", type=int, default=40, help="converged steps recorded per training stent")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--exec", choices=["graph", "eager"], default="eager",
                    help="Newton iterations launched one by one (default) or as HIP graph replays")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in launch shapes")
    ap.add_argument("--workload", choices=["stent", "sweep", "cantilever", "ring", "net"], default="stent",
                    help="stent: the braided stent (default); sweep: the stent design sweep; "
                         "cantilever, ring, net: the smaller structures, plain FP64 and FP32")
    ap.add_argument("--compare-steps", type=int, default=10,
                    help="steps of the FP32-vs-FP64-restatement comparison of the smaller structures")
    ap.add_argument("--instances", type=int, default=512, help="stents in the design sweep")
    ap.add_argument("--sweep-sdf-n", type=int, default=128)
    ap.add_argument("--sub-batch", type=int, default=0, help="instances per context in the sweep (0: autotuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device to device 5 times (scales HBM counters of a profiler run)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "copies": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(independent structures would shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "sweep":
        sweep(args)
        return
    if args.workload in ("cantilever", "ring", "net"):
        small_config(args)
        return

    from endobeams_amd import HipBatchEngine, HipEngine, abi

    from endobeams_amd import learned as L

    plain = f64b = fast = lrn = None
    batch_tuning = None
    xo = {"graph": args.exec == "graph", "autotune": not args.no_autotune}
    log = lambda name, f: print(f"# {name}: {f['steps']} steps, {f['ms_per_step']:.2f} ms/step, {f['value']:.0f} "
                                f"element-steps/s, {f['newton_iters_per_step']:.2f} Newton it./step",
                                file=sys.stderr, flush=True)
    forms = set(args.only.split(",")) if args.only else {"plain", "fp64_batch", "fast", "learned"}
    if "plain" in forms:
        conf, params, inter = stent(sdf_n=args.sdf_n)
        eng = HipEngine(conf, params, inter)
        plain = timed(eng, params, len(conf.beams), args.warmup, args.steps, **xo)
        log("plain FP64, one stent", plain)
        eng.close()
    inst = None
    if forms & {"fp64_batch", "fast", "learned"}:
        t = time.perf_counter()
        cands = [int(c) for c in args.batch_candidates.split(",")] if args.batch == 0 else [args.batch]
        inst = [stent(seed=None if s == 0 else s, sdf_n=args.sdf_n) for s in range(max(cands))]
        print(f"# {len(inst)} seeded stents built in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
        if args.batch == 0:
            from endobeams_amd import autotune
            t = time.perf_counter()
            size, rates, shapes = autotune.pick_batch_size(
                inst, cands, warmup=args.warmup, steps=3,
                log=lambda c, r: print(f"#   batch of {c}: {r:.0f} element-steps/s", file=sys.stderr, flush=True))
            batch_tuning = {"candidates": {int(k): v for k, v in rates.items()}, "chosen": size,
                            "tuning_s": time.perf_counter() - t}
            args.batch = size
            inst = inst[:size]
        print(f"# batch of {args.batch} stents", file=sys.stderr, flush=True)
    if "fp64_batch" in forms:
        eng = HipBatchEngine(inst, precision=abi.EB_PRECISION_FP64)
        f64b = timed_batch(eng, args.warmup, args.steps, **xo)
        log("FP64 batch", f64b)
        eng.close()
    if "fast" in forms:
        eng = HipBatchEngine(inst, precision=abi.EB_PRECISION_FP32)
        fast = timed_batch(eng, args.warmup, args.steps, **xo)
        log("FP32 batch", fast)
        eng.close()
    training = None
    if "learned" in forms:
        # training trajectories: 8 seeded stents not in the timed batch, FP32 form, reference predictor
        t = time.perf_counter()
        tr = [stent(seed=1000 + s, sdf_n=args.sdf_n) for s in range(8)]
        teng = HipBatchEngine(tr, precision=abi.EB_PRECISION_FP32)
        X, T = L.collect(teng, tr[0][1], args.train_steps)
        teng.close()
        t_collect = time.perf_counter() - t
        t = time.perf_counter()
        ws = L.WarmStart.train(X, T, epochs=args.epochs, batch=4096, lr=1e-3, seed=0)
        t_train = time.perf_counter() - t
        training = {"samples": int(X.shape[0]), "epochs": args.epochs, "collect_s": t_collect, "train_s": t_train,
                    "loss_first_epoch": float(ws.losses[0]), "loss_last_epoch": float(ws.losses[-1])}
        print(f"# network trained: {X.shape[0]} samples, {t_train:.1f} s, loss {ws.losses[0]:.3f} -> {ws.losses[-1]:.3f}",
              file=sys.stderr, flush=True)
        eng = HipBatchEngine(inst, precision=abi.EB_PRECISION_FP32)
        ws.install(eng)
        lrn = timed_batch(eng, args.warmup, args.steps, **xo)
        ch = L.predictor_stats(eng)
        training["network_start_kept"] = int(ch[0])
        training["fallback_to_reference_predictor"] = int(ch[1])
        training["network_skipped_by_backoff"] = int(ch[2])
        log(f"learned + FP32 batch (choices {ch.tolist()})", lrn)
        eng.close()

    # the line's value: the FP32 batch without the network when it ran. The learned warm start is
    # trained and timed in the same run and reported beside (learned_fp32_batch): on this workload its
    # start never beats the reference predictor's, so it only costs its evaluations
    main_form = next(f for f in (fast, lrn, f64b, plain) if f is not None)
    reduced = main_form is fast or main_form is lrn
    traffic = None
    if reduced and args.sdf_n == 256:
        # HBM bytes of one element assembly, from the committed FETCH_SIZE / WRITE_SIZE profiler
        # runs on 32 of these stents (scaled by a device copy of known size; 2 400 tiles: the two
        # launches of the precision form), per beam × beams
        tf = os.path.join(REPO, "profiles", "braided_stent_batch32_fp32_hbm_traffic.json")
        if os.path.exists(tf):
            ks = json.load(open(tf))["kernels"]
            names = ["ebd::" + n.strip() for n in element_kernels(True, True, 32 * 4800).split("+")]
            if all(n in ks for n in names):
                traffic = sum(ks[n]["fetch_bytes"] + ks[n]["write_bytes"] for n in names) / (32 * 4800) * \
                    active_per_launch(main_form, "beam", 0)
    avg_s = kernel_avg_s(main_form, "beam")
    peak = PEAK_FP32_TFLOPS if reduced else PEAK_FP64_TFLOPS
    s_bytes = 4 if reduced else 8
    out = {
        "metric": METRIC,
        "value": main_form["value"],
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": main_form["steps"],
        "warmup": args.warmup,
        "ms_per_step": main_form["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if reduced else "f64",
        "data": "synthetic",
        "config": {"workload": "braided stent: 24 helical wires x 200 beams (4800 beams, 4824 nodes), "
                               f"{args.sdf_n}^3 FP32 vessel SDF, Newton tolerance 1e-8",
                   "form": ("learned warm start + " if main_form is lrn else "") +
                           (f"FP32 element matrices + FP64 forces/solver, batch of {args.batch} seeded stents"
                            if reduced else "plain FP64") +
                           (", each stent with its own time step" if main_form is not plain else ", one stent") +
                           ("; the learned warm start (learned_fp32_batch) timed beside it"
                            if main_form is fast and lrn is not None else ""),
                   "dt": 1e-3, "tolerance": 1e-8, "beams_per_step": main_form["E"]},
        "newton_iters_per_step": main_form["newton_iters_per_step"],
        "newton_iters_per_step_max": main_form["newton_iters_per_step_max"],
        "failed_steps": main_form["failed_steps"],
        "wall_s_per_simulated_s": main_form["wall_s_per_simulated_s"],
        "exec": main_form["exec"],
        "roofline": {**roofline(element_kernels(reduced, True, main_form["E"]), avg_s,
                                active_per_launch(main_form, "beam", 0), peak, traffic),
                     "launches": main_form["kernels"]["beam"][1], "beams_all_instances": main_form["E"]},
        "hbm_rooflines": hbm_records(main_form, s_bytes),
        "ms_per_step_by_kernel": breakdown(main_form),
    }
    if plain is not None:
        out["plain_fp64"] = summary(plain)
        out["speedup_vs_plain_fp64"] = main_form["value"] / plain["value"]
    if f64b is not None and main_form is not f64b:
        out["fp64_batch"] = {**summary(f64b), "hbm_rooflines": hbm_records(f64b, 8)}
        out["speedup_vs_fp64_batch"] = main_form["value"] / f64b["value"]
    if fast is not None and main_form is not fast:
        out["fp32_batch_without_network"] = summary(fast)
    if lrn is not None and main_form is not lrn:
        out["learned_fp32_batch"] = summary(lrn)
        if plain is not None:
            out["learned_fp32_batch"]["speedup_vs_plain_fp64"] = lrn["value"] / plain["value"]
    if training is not None:
        out["learned_training"] = training
    if batch_tuning is not None:
        out["batch_autotune"] = batch_tuning
    if not args.no_cpu:
        cb, oracle_eng, n_cpu = cpu_baseline(args.cpu_seconds)
        out["cpu_baseline"] = cb
        # precision form against the FP64 CPU restatement after the same number of steps
        conf, params, inter = stent(sdf_n=args.sdf_n)
        g = HipEngine(conf, params, inter, precision=abi.EB_PRECISION_FP32)
        from endobeams_amd import Stepper
        st = Stepper(g, params)
        st.initialize()
        run_steps(st, n_cpu)
        err = {}
        for f in ("U_N", "R_N"):
            a, b = g.get(f), oracle_eng.get(f)
            err[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        out["fp32_error_vs_fp64_oracle"] = {"steps": n_cpu, **err}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
