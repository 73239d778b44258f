#!/usr/bin/env python3
"""Benchmark: element-steps/s and Newton iterations/step of the device Newton iteration.

One "step" is one HHT-α time step of the workload (predictor + Newton corrector to tolerance
1e-8, with the reference's adaptive Δt), inputs resident on the GPU. The default workload is the
braided stent: 24 helical wires × 200 beams (4800 beams, 4824 nodes, 28 944 dofs) crimped to
1.5 mm and released in a rigid vessel given as a 256³ FP32 signed distance field.

Prints one JSON line with the metric, the roofline of the element kernel (FP64 flops from the
op-counting restatement ÷ average launch time measured with HIP events in this run) and the CPU
restatement's throughput on a bounded sample of the same workload (host cores, 1 thread).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "endobeams.jl_amd"))

METRIC = "element-steps/s & Newton iters/step, learned+FP32 vs plain FP64 GPU, same tol"

# FP64 operations (+, -, ×, ÷; transcendentals and sqrt excluded) of one compute_beams call with
# 3 Gauss points and no damping, counted by the op-counting build of the CPU restatement
# (tests/test_flop_count.py keeps this figure in sync).
FLOPS_PER_BEAM = 26547
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def workload(name, sdf_n):
    from endobeams_amd import configs

    if name == "braided_stent":
        conf, params, inter = configs.braided_stent(sdf_n=sdf_n)
        desc = f"braided stent: 24 helical wires x 200 beams (4800 beams, 4824 nodes), {sdf_n}^3 FP32 vessel SDF"
    elif name == "net_on_sphere":
        conf, params, inter = configs.net_on_sphere()
        desc = "net of 40x40 wire crossings (3120 beams, 1600 nodes) on a 128^3 FP32 sphere SDF"
    elif name == "ring_drop":
        conf, params, inter = configs.ring_drop()
        desc = "ring of 60 beams dropped on a 64x64x32 FP32 plane SDF"
    elif name == "l_cantilever":
        conf, params, inter = configs.l_cantilever()
        desc = "L-shaped cantilever, 20 beams, tip load ramp"
    else:
        raise SystemExit(f"unknown workload {name}")
    return conf, params, inter, desc


def run_steps(stepper, n):
    """Advances until n more steps have converged; returns (converged, attempts)."""
    done = attempts = 0
    while done < n and not stepper.finished:
        attempts += 1
        if stepper.advance():
            done += 1
    return done, attempts


def cpu_baseline(args):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from endobeams_amd import Stepper

    conf, params, inter, _ = workload(args.workload, args.sdf_n)
    eng = O.OracleEngine(conf, params, inter)
    st = Stepper(eng, params)
    st.initialize()
    run_steps(st, 1)  # same first step as the GPU warm-up
    t0 = time.perf_counter()
    n0, it0 = st.stats.steps, st.stats.newton_iterations
    while time.perf_counter() - t0 < args.cpu_seconds and not st.finished:
        run_steps(st, 1)
    wall = time.perf_counter() - t0
    steps = st.stats.steps - n0
    E = len(conf.beams)
    return {"value": E * steps / wall, "unit": "element-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} converged steps of the same workload after its first step, "
                      f"{st.stats.newton_iterations - it0} Newton iterations, {wall:.1f} s, "
                      f"C++ FP64 restatement with banded-LU direct solve, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="braided_stent")
    ap.add_argument("--sdf-n", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(independent structures would shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    from endobeams_amd import HipEngine, Stepper, abi

    conf, params, inter, desc = workload(args.workload, args.sdf_n)
    prec = abi.EB_PRECISION_FP32 if args.precision == "fp32" else abi.EB_PRECISION_FP64
    eng = HipEngine(conf, params, inter, precision=prec)
    st = Stepper(eng, params)
    st.initialize()
    run_steps(st, args.warmup)
    print(f"# warm-up done: {st.stats.steps} steps", file=sys.stderr, flush=True)

    eng.profile_enable(["beam"])
    n0, it0, t_sim0, fail0 = st.stats.steps, st.stats.newton_iterations, st.tn, st.stats.failed_steps
    t0 = time.perf_counter()
    steps, attempts = run_steps(st, args.steps)
    wall = time.perf_counter() - t0
    newton = st.stats.newton_iterations - it0
    sim_t = st.tn - t_sim0
    beam_ms, beam_n = eng.profile_read("beam")
    print(f"# timed: {steps} steps, {newton} Newton iterations, {wall:.3f} s", file=sys.stderr, flush=True)

    E = len(conf.beams)
    value = E * steps / wall
    avg_s = beam_ms * 1e-3 / max(beam_n, 1)
    achieved = FLOPS_PER_BEAM * E / avg_s / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == abi.EB_PRECISION_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": desc, "form": "plain FP64" if prec == abi.EB_PRECISION_FP64 else "FP32 element/contact",
                   "dt": params.initial_timestep, "tolerance": params.tolerance_residual,
                   "beams": E, "dofs": 6 * len(conf.nodes)},
        "newton_iters_per_step": newton / max(steps, 1),
        "failed_steps": st.stats.failed_steps - fail0,
        "wall_s_per_simulated_s": wall / sim_t if sim_t > 0 else None,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "beam_kernel<double>", "avg_launch_us": avg_s * 1e6, "launches": beam_n,
                     "flops_per_launch": FLOPS_PER_BEAM * E},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
