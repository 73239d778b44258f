"""ctypes mirror of include/endobeams_hip.h (the C ABI of libendobeams_hip.so).

Only declarations live here: structs, constants and the prototype table shared by every loader
of a library that exports this ABI.
"""
import ctypes as C

EB_ABI_VERSION = 1
EB_NPROPS = 14

EB_OK, EB_ERR_ARG, EB_ERR_DEVICE, EB_ERR_SOLVE, EB_ERR_STATE = 0, -1, -2, -3, -4
# device staging of eb_output_points_noalloc, in doubles (include/endobeams_hip.h)
EB_NOALLOC_SCRATCH_DOUBLES = 1 << 22

EB_SURFACE_NONE, EB_SURFACE_SPHERE, EB_SURFACE_SDF = 0, 1, 2
EB_PRECISION_FP64, EB_PRECISION_FP32 = 0, 1

FIELDS = [
    "U", "UDOT", "UDDOT", "W", "WDOT", "WDDOT", "R", "DR",
    "U_N", "UDOT_N", "UDDOT_N", "W_N", "WDOT_N", "WDDOT_N", "R_N", "DR_N",
    "FEXT", "TINT", "TK", "TC", "FEXT_N", "TINT_N", "TK_N", "TC_N",
    "D", "DDOT", "DDDOT", "RES", "DELTA_D",
    "K", "C", "M", "K_N", "C_N", "M_N", "KTAN",
    "RES_FREE", "KTAN_FREE", "DELTA_D_FREE", "IMPOSED", "ENERGY",
]
F = {name: i for i, name in enumerate(FIELDS)}

NODE_VEC_FIELDS = {"U", "UDOT", "UDDOT", "W", "WDOT", "WDDOT", "U_N", "UDOT_N", "UDDOT_N", "W_N", "WDOT_N", "WDDOT_N"}
NODE_MAT_FIELDS = {"R", "DR", "R_N", "DR_N"}
DOF_FIELDS = {"FEXT", "TINT", "TK", "TC", "FEXT_N", "TINT_N", "TK_N", "TC_N", "D", "DDOT", "DDDOT", "RES", "DELTA_D"}
NNZ_FIELDS = {"K", "C", "M", "K_N", "C_N", "M_N", "KTAN"}

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_int64)


class ModelDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("precision", C.c_int32),
        ("n_nodes", C.c_int64), ("n_beams", C.c_int64),
        ("X0", _dp), ("conn", _lp), ("l0", _dp), ("Re0", _dp), ("props", _dp),
        ("n_fixed", C.c_int64), ("fixed_dofs", _lp),
        ("n_gauss", C.c_int32), ("pad0", C.c_int32),
        ("gauss_w", _dp), ("gauss_z", _dp),
        ("surface", C.c_int32), ("sdf_prefiltered", C.c_int32),
        ("kn", C.c_double), ("mu", C.c_double), ("eps_t", C.c_double), ("eta_n", C.c_double),
        ("n_contact_beams", C.c_int64), ("contact_beams", _lp),
        ("sphere_center", C.c_double * 3), ("sphere_radius", C.c_double),
        ("sdf_n", C.c_int64 * 3), ("sdf_d", C.c_double * 3), ("sdf_x0", C.c_double * 3),
        ("sdf_samples", _dp),
    ]


class NewtonParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
        ("tol_res", C.c_double), ("tol_disp", C.c_double),
        ("max_iterations", C.c_int32), ("krylov_maxit", C.c_int32),
        ("krylov_tol", C.c_double),
    ]


class TimeParams(C.Structure):
    _fields_ = [
        ("initial_timestep", C.c_double), ("min_timestep", C.c_double), ("max_timestep", C.c_double),
        ("simulation_end_time", C.c_double), ("accelerate_after_success_it", C.c_int32),
        ("stop_on_energy_threshold", C.c_int32), ("energy_threshold", C.c_double),
    ]


class RunState(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double), ("successive", C.c_int32), ("stopped", C.c_int32),
        ("steps", C.c_int64), ("failed_steps", C.c_int64), ("newton_iterations", C.c_int64),
        ("last_k", C.c_int32), ("pad", C.c_int32),
    ]


# int (*)(eb_ctx*, double t, void* user)
STEP_CALLBACK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_void_p)
# int (*)(eb_ctx*, int64_t instance, double t, void* user)
INSTANCE_CALLBACK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_double, C.c_void_p)


class Sizes(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("n_nodes", "n_beams", "ndofs", "nnz", "n_free", "n_fixed", "nnz_free")]


_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_NP = C.POINTER(NewtonParams)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes); the context argument is a void* here
PROTOTYPES = {
    "create": (_i, [C.POINTER(ModelDesc), C.POINTER(_vp)]),
    "destroy": (_i, [_vp]),
    "last_error": (C.c_char_p, [_vp]),
    "get_sizes": (_i, [_vp, C.POINTER(Sizes)]),
    "get_pattern": (_i, [_vp, _lp, _lp, _lp, _lp]),
    "get_field": (_i, [_vp, _i, _dp, C.c_int64]),
    "set_field": (_i, [_vp, _i, _dp, C.c_int64]),
    "assemble": (_i, [_vp]),
    "assemble_contact": (_i, [_vp, _d]),
    "tangent_residuals_corrector": (_i, [_vp, _d, _d, _d, _d]),
    "tangent_residuals_predictor": (_i, [_vp, _d, _d, _d, _d]),
    "apply_boundary_conditions": (_i, [_vp]),
    "extract_free_dofs": (_i, [_vp]),
    "solve_free_dofs": (_i, [_vp, _NP, _IP]),
    "initialize_global_corrector": (_i, [_vp]),
    "initialize_global_predictor": (_i, [_vp]),
    "compute_global_corrector": (_i, [_vp, _d, _d, _d]),
    "compute_global_predictor": (_i, [_vp, _d, _d, _d]),
    "compute_local_corrector": (_i, [_vp, _d, _d, _d]),
    "compute_local_predictor": (_i, [_vp, _d, _d, _d]),
    "compute_norms_corrector": (_i, [_vp, _dp, _dp]),
    "update_loads": (_i, [_vp, C.c_int64, _lp, _dp]),
    "update_boundary_conditions": (_i, [_vp, _dp, C.c_int64]),
    "initialize_state_simulation": (_i, [_vp]),
    "predictor": (_i, [_vp, _NP, _d]),
    "corrector": (_i, [_vp, _NP, _d, _IP]),
    "solve_step_dynamics": (_i, [_vp, _NP, _d, _IP]),
    "update_converged": (_i, [_vp]),
    "update_not_converged": (_i, [_vp]),
    "run_steps": (_i, [_vp, _NP, C.POINTER(TimeParams), C.c_int64, STEP_CALLBACK, _vp, C.POINTER(RunState)]),
    "run_steps_batch": (_i, [_vp, _NP, C.POINTER(TimeParams), C.c_int64, INSTANCE_CALLBACK, _vp, C.POINTER(RunState)]),
    "get_counters": (_i, [_vp, _lp]),
    "profile_enable": (_i, [_vp, C.c_uint32]),
    "profile_read": (_i, [_vp, _i, _dp, _lp]),
    "create_batch": (_i, [C.POINTER(ModelDesc), C.c_int64, C.POINTER(_vp)]),
    "get_instances": (_i, [_vp, _lp, _lp, _lp]),
    "get_instance_iterations": (_i, [_vp, C.POINTER(C.c_int32)]),
    "get_instance_norms": (_i, [_vp, _dp]),
    "output_points": (_i, [_vp, C.c_int32, _dp, _dp, _dp, _dp, _dp]),
    "output_points_noalloc": (_i, [_vp, C.c_int32, _dp, _dp, _dp, _dp, _dp]),
    "get_contact_stats": (_i, [_vp, _lp]),
    "get_work_counters": (_i, [_vp, _lp]),
    "get_sdf_coefficients": (_i, [_vp, C.c_int64, _dp, C.c_int64]),
    "set_exec_mode": (_i, [_vp, _i]),
    "get_exec_stats": (_i, [_vp, C.POINTER(C.c_int32), _lp, _lp]),
    "autotune": (_i, [_vp, C.POINTER(C.c_int32), _dp]),
    "set_tuning": (_i, [_vp, C.POINTER(C.c_int32)]),
    "get_tuning": (_i, [_vp, C.POINTER(C.c_int32)]),
}

EB_EXEC_EAGER, EB_EXEC_GRAPH, EB_EXEC_FUSED = 0, 1, 2
TUNE_NAMES = ("chain_lds_kb", "chain_solve_lds", "gather_block", "chain_algo")
EB_TUNE_COUNT, EB_TUNE_CANDIDATES = 4, 4
EB_CHAIN_CYCLIC_REDUCTION, EB_CHAIN_BLOCK_LU = 1, 2
TUNE_CANDIDATES = ((16, 32, 48, 96), (1, 2), (64, 128, 256), (EB_CHAIN_CYCLIC_REDUCTION, EB_CHAIN_BLOCK_LU))

KERNELS = {"beam": 0, "contact": 1, "gather": 2, "precond": 3, "krylov": 4, "update": 5, "gather_matrices": 6}

# include/endobeams_learned.h (network handle is a void* here)
_fp = C.POINTER(C.c_float)
LEARNED_PROTOTYPES = {
    "mlp_create": (_i, [C.POINTER(C.c_int32), C.c_int32, C.c_uint64, C.POINTER(_vp)]),
    "mlp_destroy": (_i, [_vp]),
    "mlp_last_error": (C.c_char_p, [_vp]),
    "mlp_num_params": (_i, [_vp, _lp]),
    "mlp_get_params": (_i, [_vp, _fp, C.c_int64]),
    "mlp_set_params": (_i, [_vp, _fp, C.c_int64]),
    "mlp_forward": (_i, [_vp, _fp, C.c_int64, _fp]),
    "mlp_loss_gradient": (_i, [_vp, _fp, _fp, C.c_int64, _fp, _fp]),
    "mlp_set_dataset": (_i, [_vp, _fp, _fp, C.c_int64]),
    "mlp_train": (_i, [_vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp]),
    "learned_sample": (_i, [_vp, _fp, _fp]),
    "set_learned_predictor": (_i, [_vp, _vp, _fp, _fp]),
    "set_learned_backoff": (_i, [_vp, C.c_int32]),
    "get_predictor_stats": (_i, [_vp, _lp]),
    "get_learned_increment": (_i, [_vp, _dp, C.c_int64]),
}
LEARNED_INPUTS, LEARNED_OUTPUTS = 27, 6


def bind(lib, prefix):
    """Attach restype/argtypes for every ABI function found in `lib` under `prefix`; returns
    {short_name: function}. Raises if any declared function is missing."""
    out = {}
    protos = dict(PROTOTYPES)
    if hasattr(lib, prefix + "mlp_create"):
        protos.update(LEARNED_PROTOTYPES)
    for name, (res, args) in protos.items():
        fn = getattr(lib, prefix + name)  # AttributeError if the symbol is absent
        fn.restype = res
        fn.argtypes = args
        out[name] = fn
    return out
