"""Python handle on a library that exports the EndoBeams C ABI (include/endobeams_hip.h).

`Engine` wraps one eb_ctx: it packs the model, uploads the initial node state and Forces(conf)
(src/def_structures.jl:10-27), and exposes one method per ABI entry point with the reference's
function names. `HipEngine` is the product: it loads libendobeams_hip.so and refuses to run
without it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import abi
from .model import BeamsConfiguration, ModelArrays, RigidInteraction, SimulationParams, initial_fields

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "lib", "libendobeams_hip.so")


class EBError(RuntimeError):
    pass


class Engine:
    prefix = "eb_"

    def __init__(self, lib: C.CDLL, conf: BeamsConfiguration, params: SimulationParams,
                 inter: Optional[RigidInteraction] = None, precision: int = abi.EB_PRECISION_FP64):
        self.lib = lib
        self.fn = abi.bind(lib, self.prefix)
        self.conf, self.params, self.inter = conf, params, inter
        self.arrays = ModelArrays(conf, params, inter, precision)
        h = C.c_void_p()
        st = self.fn["create"](C.byref(self.arrays.desc), C.byref(h))
        if st != abi.EB_OK:
            raise EBError(f"{self.prefix}create failed with status {st}")
        self.h = h
        s = abi.Sizes()
        self._check(self.fn["get_sizes"](self.h, C.byref(s)), "get_sizes")
        self.sizes = s
        for name, arr in initial_fields(conf).items():
            self.set(name, arr)
        # Forces(conf): fᵉˣᵗ[i] = f(0, i) in both forcesⁿ and forcesⁿ⁺¹
        ld = conf.loads
        if ld is not None and ld.concentrated_force is not None:
            f0 = np.zeros(conf.ndofs)
            dofs = np.asarray(ld.concentrated_force.loaded_dofs, dtype=np.int64)
            f0[dofs - 1] = ld.concentrated_force.values(0.0)
            self.set("FEXT", f0)
            self.set("FEXT_N", f0)
        self._np = params.newton()

    # -- plumbing -------------------------------------------------------------------------------
    def _check(self, st, what):
        if st != abi.EB_OK:
            msg = self.fn["last_error"](self.h)
            raise EBError(f"{self.prefix}{what} failed ({st}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.fn["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field_len(self, name: str) -> int:
        s = self.sizes
        if name in abi.NODE_VEC_FIELDS:
            return 3 * s.n_nodes
        if name in abi.NODE_MAT_FIELDS:
            return 9 * s.n_nodes
        if name in abi.DOF_FIELDS:
            return s.ndofs
        if name in abi.NNZ_FIELDS:
            return s.nnz
        return {"RES_FREE": s.n_free, "DELTA_D_FREE": s.n_free, "KTAN_FREE": s.nnz_free,
                "IMPOSED": s.n_fixed, "ENERGY": 3}[name]

    def get(self, name: str) -> np.ndarray:
        n = self.field_len(name)
        out = np.empty(n, dtype=np.float64)
        self._check(self.fn["get_field"](self.h, abi.F[name], out.ctypes.data_as(C.POINTER(C.c_double)), n), f"get_field({name})")
        return out

    def set(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self._check(self.fn["set_field"](self.h, abi.F[name], a.ctypes.data_as(C.POINTER(C.c_double)), a.size), f"set_field({name})")

    def pattern(self):
        s = self.sizes
        colptr = np.empty(s.ndofs + 1, np.int64)
        rowval = np.empty(s.nnz, np.int64)
        smap = np.empty(144 * s.n_beams, np.int64)
        sfree = np.empty(max(s.nnz_free, 1), np.int64)
        lp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        self._check(self.fn["get_pattern"](self.h, lp(colptr), lp(rowval), lp(smap), lp(sfree)), "get_pattern")
        return colptr, rowval, smap.reshape(s.n_beams, 144), sfree[: s.nnz_free]

    def counters(self):
        out = np.zeros(4, np.int64)
        self._check(self.fn["get_counters"](self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "get_counters")
        return out

    def contact_stats(self):
        """(beams listed in contact, their Gauss points in contact), summed over the contact-tangent
        launches since the context was created."""
        out = np.zeros(2, np.int64)
        self._check(self.fn["get_contact_stats"](self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "get_contact_stats")
        return int(out[0]), int(out[1])

    def work_counters(self):
        """(beams through the element kernel, tangent entries through the matrix gather, chain
        positions factored) over the launches timed since profile_enable, counting only the
        instances active in each launch."""
        out = np.zeros(3, np.int64)
        self._check(self.fn["get_work_counters"](self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "get_work_counters")
        return int(out[0]), int(out[1]), int(out[2])

    def profile_enable(self, kernels=()):
        mask = 0
        for k in kernels:
            mask |= 1 << abi.KERNELS[k]
        self._check(self.fn["profile_enable"](self.h, mask), "profile_enable")

    def profile_read(self, kernel):
        ms, n = C.c_double(0), C.c_int64(0)
        self._check(self.fn["profile_read"](self.h, abi.KERNELS[kernel], C.byref(ms), C.byref(n)), "profile_read")
        return ms.value, n.value

    def output_points(self, n_points: int = 10) -> dict:
        """recompute_at_gausspts_beams! (src/IO/visualization_beams.jl:132-181): position,
        displacement, velocity (n_beams·n_points × 3), strain and stress (n_beams·n_points) at
        n_points points along every beam, from the current node state."""
        m = self.sizes.n_beams * n_points
        out = {"position": np.empty((m, 3)), "displacement": np.empty((m, 3)), "velocity": np.empty((m, 3)),
               "strain": np.empty(m), "stress": np.empty(m)}
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self.fn["output_points"](self.h, int(n_points), dp(out["position"]), dp(out["displacement"]),
                                             dp(out["velocity"]), dp(out["strain"]), dp(out["stress"])), "output_points")
        return out

    def sdf_coefficients(self, instance: int = 0) -> np.ndarray:
        """The padded quadratic B-spline coefficients of an instance's SDF as the engine holds them
        ((n2+2, n1+2, n0+2), x fastest)."""
        conf, _, inter = self.instances[instance] if hasattr(self, "instances") else (self.conf, None, self.inter)
        n = inter.master.n
        out = np.empty((n[2] + 2) * (n[1] + 2) * (n[0] + 2))
        self._check(self.fn["get_sdf_coefficients"](self.h, instance, out.ctypes.data_as(C.POINTER(C.c_double)), out.size),
                    "get_sdf_coefficients")
        return out.reshape(n[2] + 2, n[1] + 2, n[0] + 2)

    # -- execution mode and launch shapes (no reference counterpart) --------------------------------
    def set_exec_mode(self, graph):
        """Newton iterations as replays of a captured HIP graph (True) or launched one by one
        (False); "fused": each time step (predictor and every Newton iteration) as one launch of
        one workgroup, for one small structure whose chain preconditioner holds every coupling."""
        if graph == "fused":
            mode = abi.EB_EXEC_FUSED
        else:
            mode = abi.EB_EXEC_GRAPH if graph else abi.EB_EXEC_EAGER
        self._check(self.fn["set_exec_mode"](self.h, mode), "set_exec_mode")

    def exec_stats(self):
        """(mode, graph captures, graph replays)"""
        m, c, r = C.c_int32(0), C.c_int64(0), C.c_int64(0)
        self._check(self.fn["get_exec_stats"](self.h, C.byref(m), C.byref(c), C.byref(r)), "get_exec_stats")
        return m.value, c.value, r.value

    def autotune(self):
        """Times the candidate launch shapes on this context and keeps the fastest; returns
        ({name: choice}, {name: {candidate: ms}})."""
        ch = np.zeros(abi.EB_TUNE_COUNT, np.int32)
        ms = np.zeros(abi.EB_TUNE_COUNT * abi.EB_TUNE_CANDIDATES, np.float64)
        self._check(self.fn["autotune"](self.h, ch.ctypes.data_as(C.POINTER(C.c_int32)),
                                        ms.ctypes.data_as(C.POINTER(C.c_double))), "autotune")
        times = {}
        for i, name in enumerate(abi.TUNE_NAMES):
            times[name] = {c: float(ms[i * abi.EB_TUNE_CANDIDATES + j])
                           for j, c in enumerate(abi.TUNE_CANDIDATES[i]) if ms[i * abi.EB_TUNE_CANDIDATES + j] >= 0}
        return dict(zip(abi.TUNE_NAMES, (int(v) for v in ch))), times

    @property
    def tuning(self):
        ch = np.zeros(abi.EB_TUNE_COUNT, np.int32)
        self._check(self.fn["get_tuning"](self.h, ch.ctypes.data_as(C.POINTER(C.c_int32))), "get_tuning")
        return dict(zip(abi.TUNE_NAMES, (int(v) for v in ch)))

    @tuning.setter
    def tuning(self, choice):
        ch = np.array([int(choice.get(n, 0)) for n in abi.TUNE_NAMES], np.int32)
        self._check(self.fn["set_tuning"](self.h, ch.ctypes.data_as(C.POINTER(C.c_int32))), "set_tuning")

    # -- reference functions ----------------------------------------------------------------------
    def assemble(self):
        self._check(self.fn["assemble"](self.h), "assemble")

    def assemble_contact(self, alpha):
        self._check(self.fn["assemble_contact"](self.h, alpha), "assemble_contact")

    def compute_tangent_and_residuals_corrector(self, dt, a, b, g):
        self._check(self.fn["tangent_residuals_corrector"](self.h, dt, a, b, g), "tangent_residuals_corrector")

    def compute_tangent_and_residuals_predictor(self, dt, a, b, g):
        self._check(self.fn["tangent_residuals_predictor"](self.h, dt, a, b, g), "tangent_residuals_predictor")

    def apply_boundary_conditions(self):
        self._check(self.fn["apply_boundary_conditions"](self.h), "apply_boundary_conditions")

    def extract_free_dofs(self):
        self._check(self.fn["extract_free_dofs"](self.h), "extract_free_dofs")

    def solve_free_dofs(self):
        its = C.c_int(0)
        self._check(self.fn["solve_free_dofs"](self.h, C.byref(self._np), C.byref(its)), "solve_free_dofs")
        return its.value

    def initialize_global_corrector(self):
        self._check(self.fn["initialize_global_corrector"](self.h), "initialize_global_corrector")

    def initialize_global_predictor(self):
        self._check(self.fn["initialize_global_predictor"](self.h), "initialize_global_predictor")

    def compute_global_corrector(self, dt, b, g):
        self._check(self.fn["compute_global_corrector"](self.h, dt, b, g), "compute_global_corrector")

    def compute_global_predictor(self, dt, b, g):
        self._check(self.fn["compute_global_predictor"](self.h, dt, b, g), "compute_global_predictor")

    def compute_local_corrector(self, dt, b, g):
        self._check(self.fn["compute_local_corrector"](self.h, dt, b, g), "compute_local_corrector")

    def compute_local_predictor(self, dt, b, g):
        self._check(self.fn["compute_local_predictor"](self.h, dt, b, g), "compute_local_predictor")

    def compute_norms_corrector(self):
        r, d = C.c_double(0), C.c_double(0)
        self._check(self.fn["compute_norms_corrector"](self.h, C.byref(r), C.byref(d)), "compute_norms_corrector")
        return r.value, d.value

    def update_loads(self, t):
        ld = self.conf.loads
        if ld is None or ld.concentrated_force is None:
            return
        cf = ld.concentrated_force
        dofs = np.ascontiguousarray(cf.loaded_dofs, dtype=np.int64)
        vals = cf.values(t)
        self._check(self.fn["update_loads"](self.h, dofs.size, dofs.ctypes.data_as(C.POINTER(C.c_int64)),
                                            vals.ctypes.data_as(C.POINTER(C.c_double))), "update_loads")

    def update_boundary_conditions(self, t):
        bcs = self.conf.bcs
        if bcs.imposed_displacements_function is None or len(bcs.displaced_indices) == 0:
            return
        for i in bcs.displaced_indices:
            bcs.imposed_displacements[i] = bcs.imposed_displacements_function(t)
        v = np.ascontiguousarray(bcs.imposed_displacements, dtype=np.float64)
        self._check(self.fn["update_boundary_conditions"](self.h, v.ctypes.data_as(C.POINTER(C.c_double)), v.size),
                    "update_boundary_conditions")

    def initialize_state_simulation(self):
        self._check(self.fn["initialize_state_simulation"](self.h), "initialize_state_simulation")

    def predictor(self, dt):
        self._check(self.fn["predictor"](self.h, C.byref(self._np), dt), "predictor")

    def corrector(self, dt):
        k = C.c_int(0)
        self._check(self.fn["corrector"](self.h, C.byref(self._np), dt, C.byref(k)), "corrector")
        return k.value

    def solve_step_dynamics(self, t, dt):
        """solve_step_dynamics! (src/solve_step_dynamics.jl:1-14): loads and imposed displacements
        at tⁿ⁺¹, then predictor and corrector; returns the corrector's k."""
        self.update_loads(t)
        self.update_boundary_conditions(t)
        k = C.c_int(0)
        self._check(self.fn["solve_step_dynamics"](self.h, C.byref(self._np), dt, C.byref(k)), "solve_step_dynamics")
        return k.value

    def run_steps(self, max_steps: int, state: Optional[abi.RunState] = None) -> abi.RunState:
        """run_simulation!'s time loop inside the library (eb_run_steps): up to `max_steps` more
        converged steps from `state` (a fresh one: t = 0, Δt = initial_timestep). Loads and
        imposed displacements of each attempt's time are set by a callback into this engine."""
        p = self.params
        if state is None:
            state = abi.RunState(t=0.0, dt=p.initial_timestep)
        tp = abi.TimeParams(p.initial_timestep, p.min_timestep, p.max_timestep, p.simulation_end_time,
                            int(p.accelerate_after_success_it), int(bool(p.stop_on_energy_threshold)),
                            float(p.energy_threshold))
        errors = []

        def on_step(ctx, t, user):
            try:
                self.update_loads(t)
                self.update_boundary_conditions(t)
                return 0
            except Exception as e:  # reported after eb_run_steps returns
                errors.append(e)
                return 1
        cb = abi.STEP_CALLBACK(on_step)
        st = self.fn["run_steps"](self.h, C.byref(self._np), C.byref(tp), int(max_steps), cb, None, C.byref(state))
        if errors:
            raise errors[0]
        self._check(st, "run_steps")
        return state

    def _time_params(self):
        p = self.params
        return abi.TimeParams(p.initial_timestep, p.min_timestep, p.max_timestep, p.simulation_end_time,
                              int(p.accelerate_after_success_it), int(bool(p.stop_on_energy_threshold)),
                              float(p.energy_threshold))

    def _has_time_functions(self) -> bool:
        confs = [c for (c, _, _) in self.instances] if hasattr(self, "instances") else [self.conf]
        for c in confs:
            if c.loads is not None and c.loads.concentrated_force is not None:
                return True
            if c.bcs.imposed_displacements_function is not None and len(c.bcs.displaced_indices) > 0:
                return True
        return False

    def update_instance(self, s: int, t: float):
        """Loads and imposed displacements of instance s at its time t (a single structure: s = 0)."""
        self.update_loads(t)
        self.update_boundary_conditions(t)

    def run_steps_batch(self, max_steps: int, states=None):
        """eb_run_steps_batch: every instance advances with its own time and Δt (run_simulation!'s
        loop per instance), all of them stepping together until none that has not reached the end
        time or aborted (stopped = 2) has fewer than `max_steps` more converged steps (some may end
        with more). `states` is a ctypes array of RunState, one per instance (fresh ones: t = 0,
        Δt = initial_timestep); returns it."""
        S = getattr(self, "S", 1)
        if states is None:
            states = (abi.RunState * S)()
            for r in states:
                r.t, r.dt = 0.0, self.params.initial_timestep
        tp = self._time_params()
        errors = []

        def on_step(ctx, s, t, user):
            try:
                self.update_instance(int(s), t)
                return 0
            except Exception as e:  # reported after eb_run_steps_batch returns
                errors.append(e)
                return 1
        cb = abi.INSTANCE_CALLBACK(on_step) if self._has_time_functions() else abi.INSTANCE_CALLBACK()
        st = self.fn["run_steps_batch"](self.h, C.byref(self._np), C.byref(tp), int(max_steps), cb, None, states)
        if errors:
            raise errors[0]
        self._check(st, "run_steps_batch")
        return states

    def update_converged(self):
        self._check(self.fn["update_converged"](self.h), "update_converged")

    def update_not_converged(self):
        self._check(self.fn["update_not_converged"](self.h), "update_not_converged")


class BatchEngine(Engine):
    """Many independent structures in one context (eb_create_batch): design sweeps advanced
    together. Instance s keeps its own loads, boundary conditions and contact surface; the batch
    shares the time step, the HHT/Newton parameters and the contact parameters. Fields are the
    instances' fields one after the other (`offsets()` gives the node and beam offsets)."""

    def __init__(self, lib: C.CDLL, instances, precision: int = abi.EB_PRECISION_FP64):
        self.lib = lib
        self.fn = abi.bind(lib, self.prefix)
        self.instances = [(c, p, i) for (c, p, i) in instances]
        conf0, params0, inter0 = self.instances[0]
        self.conf, self.params, self.inter = conf0, params0, inter0
        self.arrays_list = [ModelArrays(c, p, i, precision) for (c, p, i) in self.instances]
        S = len(self.arrays_list)
        descs = (abi.ModelDesc * S)(*[a.desc for a in self.arrays_list])
        h = C.c_void_p()
        st = self.fn["create_batch"](descs, S, C.byref(h))
        if st != abi.EB_OK:
            raise EBError(f"{self.prefix}create_batch failed with status {st}")
        self.h = h
        self._descs = descs
        s = abi.Sizes()
        self._check(self.fn["get_sizes"](self.h, C.byref(s)), "get_sizes")
        self.sizes = s
        self.S = S
        noff = np.zeros(S + 1, np.int64)
        boff = np.zeros(S + 1, np.int64)
        n = C.c_int64(0)
        lp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        self._check(self.fn["get_instances"](self.h, C.byref(n), lp(noff), lp(boff)), "get_instances")
        self.node_off, self.beam_off = noff, boff
        fields = [initial_fields(c) for (c, _, _) in self.instances]
        for name in fields[0]:
            self.set(name, np.concatenate([np.asarray(f[name], dtype=np.float64).reshape(-1) for f in fields]))
        f0 = np.concatenate([self._forces(c, 0.0) for (c, _, _) in self.instances])
        if f0.any():
            self.set("FEXT", f0)
            self.set("FEXT_N", f0)
        self._np = params0.newton()

    @staticmethod
    def _forces(conf, t):
        f = np.zeros(conf.ndofs)
        ld = conf.loads
        if ld is not None and ld.concentrated_force is not None:
            dofs = np.asarray(ld.concentrated_force.loaded_dofs, dtype=np.int64)
            f[dofs - 1] = ld.concentrated_force.values(t)
        return f

    def offsets(self):
        return self.node_off.copy(), self.beam_off.copy()

    def instance_slice(self, name: str, s: int) -> slice:
        """Slice of field `name` that belongs to instance s."""
        a, b = self.node_off[s], self.node_off[s + 1]
        if name in abi.NODE_VEC_FIELDS:
            return slice(3 * a, 3 * b)
        if name in abi.NODE_MAT_FIELDS:
            return slice(9 * a, 9 * b)
        if name in abi.DOF_FIELDS:
            return slice(6 * a, 6 * b)
        raise KeyError(f"{name} has no per-instance slice")

    def instance_iterations(self) -> np.ndarray:
        k = np.zeros(self.S, np.int32)
        self._check(self.fn["get_instance_iterations"](self.h, k.ctypes.data_as(C.POINTER(C.c_int32))),
                    "get_instance_iterations")
        return k

    def update_loads(self, t):
        dofs, vals = [], []
        for s, (c, _, _) in enumerate(self.instances):
            ld = c.loads
            if ld is None or ld.concentrated_force is None:
                continue
            cf = ld.concentrated_force
            dofs.append(np.asarray(cf.loaded_dofs, dtype=np.int64) + 6 * int(self.node_off[s]))
            vals.append(np.asarray(cf.values(t), dtype=np.float64))
        if not dofs:
            return
        d = np.ascontiguousarray(np.concatenate(dofs))
        v = np.ascontiguousarray(np.concatenate(vals))
        self._check(self.fn["update_loads"](self.h, d.size, d.ctypes.data_as(C.POINTER(C.c_int64)),
                                            v.ctypes.data_as(C.POINTER(C.c_double))), "update_loads")

    def update_instance(self, s: int, t: float):
        """Loads and imposed displacements of instance s alone at its own time t."""
        c = self.instances[s][0]
        ld = c.loads
        if ld is not None and ld.concentrated_force is not None:
            cf = ld.concentrated_force
            d = np.ascontiguousarray(np.asarray(cf.loaded_dofs, dtype=np.int64) + 6 * int(self.node_off[s]))
            v = np.ascontiguousarray(cf.values(t), dtype=np.float64)
            self._check(self.fn["update_loads"](self.h, d.size, d.ctypes.data_as(C.POINTER(C.c_int64)),
                                                v.ctypes.data_as(C.POINTER(C.c_double))), "update_loads")
        bcs = c.bcs
        if bcs.imposed_displacements_function is not None and len(bcs.displaced_indices) > 0:
            for i in bcs.displaced_indices:
                bcs.imposed_displacements[i] = bcs.imposed_displacements_function(t)
            v = np.ascontiguousarray(np.concatenate([np.asarray(ci.bcs.imposed_displacements, dtype=np.float64).reshape(-1)
                                                     for (ci, _, _) in self.instances]))
            self._check(self.fn["update_boundary_conditions"](self.h, v.ctypes.data_as(C.POINTER(C.c_double)), v.size),
                        "update_boundary_conditions")

    def update_boundary_conditions(self, t):
        parts, any_fn = [], False
        for (c, _, _) in self.instances:
            bcs = c.bcs
            if bcs.imposed_displacements_function is not None and len(bcs.displaced_indices) > 0:
                any_fn = True
                for i in bcs.displaced_indices:
                    bcs.imposed_displacements[i] = bcs.imposed_displacements_function(t)
            parts.append(np.asarray(bcs.imposed_displacements, dtype=np.float64).reshape(-1))
        if not any_fn:
            return
        v = np.ascontiguousarray(np.concatenate(parts))
        self._check(self.fn["update_boundary_conditions"](self.h, v.ctypes.data_as(C.POINTER(C.c_double)), v.size),
                    "update_boundary_conditions")


_HIP_LIB = None
_HIP_LIB_PATH = None


def load_hip_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libendobeams_hip.so; raises if it has not been built (there is no CPU fallback).
    One build per process: two builds of the library export the same kernel and launcher
    symbols, and the dynamic linker would bind the second one's launches to the first one's
    kernels, so asking for a second path raises instead of silently running the first."""
    global _HIP_LIB, _HIP_LIB_PATH
    real = os.path.realpath(path)
    if _HIP_LIB is None:
        if not os.path.exists(path):
            raise EBError(f"{path} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
        _HIP_LIB = C.CDLL(path, mode=C.RTLD_GLOBAL)
        _HIP_LIB_PATH = real
    elif real != _HIP_LIB_PATH:
        raise EBError(f"{path}: this process already runs {_HIP_LIB_PATH}; load another build in another process")
    return _HIP_LIB


class HipEngine(Engine):
    """The MI355X path: every entry point runs HIP kernels on the current device."""
    prefix = "eb_"

    def __init__(self, conf, params, inter=None, precision=abi.EB_PRECISION_FP64, lib_path: str = LIB_PATH):
        super().__init__(load_hip_library(lib_path), conf, params, inter, precision)


class HipBatchEngine(BatchEngine):
    """Batch of independent structures on the MI355X path; `instances` is a list of
    (BeamsConfiguration, SimulationParams, RigidInteraction or None)."""
    prefix = "eb_"

    def __init__(self, instances, precision=abi.EB_PRECISION_FP64, lib_path: str = LIB_PATH):
        super().__init__(load_hip_library(lib_path), instances, precision)
