/*
 * endobeams_hip.h — C ABI of libendobeams_hip.so, the MI355X (gfx950) implementation of the
 * per-Newton-iteration hot path of EndoBeams.jl (assembly + beam-to-rigid contact + HHT-alpha
 * residual/tangent + linear solve + Newmark/rotation update).
 *
 * Conventions (all entry points):
 *   - plain C types only; every function returns int status (0 = OK, < 0 = error code) and never
 *     throws; the message of the last error is returned by eb_last_error(ctx).
 *   - the caller owns every host buffer; the library owns every device buffer.
 *   - calls are synchronous with respect to the host buffers they read or write.
 *   - a context is not thread safe: use one per host task.
 *   - index arrays are 1-based Int64, exactly as Julia stores them (node indices, dofs, beams).
 *   - vector fields use Julia's in-memory layouts: Vector{SVector{3,Float64}} is 3*N doubles
 *     node-major (x,y,z); Vector{SMatrix{3,3,Float64}} is 9*N doubles, column-major per node;
 *     SparseMatrixCSC nzval arrays are in CSC order of the pattern built by sparsity_beams.
 *
 * Each function names the reference function it replaces (EndoBeams.jl, path:line).
 */
#ifndef ENDOBEAMS_HIP_H
#define ENDOBEAMS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EB_ABI_VERSION 1

/* status codes */
#define EB_OK 0
#define EB_ERR_ARG (-1)      /* invalid argument / size mismatch */
#define EB_ERR_DEVICE (-2)   /* HIP runtime error */
#define EB_ERR_SOLVE (-3)    /* linear solve failed (breakdown / singular) */
#define EB_ERR_STATE (-4)    /* call not valid in the current state */

/* per-beam property record, 14 doubles per beam, in the order of BeamMaterialProperties
 * (src/beams/def_beam.jl:19-26, :74-85, :135-149):
 * radius, E, K̄ū, K̄Θ̅[1..3], K̄Θ̅Θ̅[1..3], Jᵨ[1..3], Aᵨ, damping */
#define EB_NPROPS 14

/* rigid master surfaces (src/interactions/def_interactions.jl:24-63) */
#define EB_SURFACE_NONE 0
#define EB_SURFACE_SPHERE 1   /* SphereSurface(center, radius) */
#define EB_SURFACE_SDF 2      /* DiscreteSignedDistanceField, quadratic B-spline Reflect(OnCell()) */

/* arithmetic of the element and contact kernels */
#define EB_PRECISION_FP64 0   /* everything FP64 (parity form) */
#define EB_PRECISION_FP32 1   /* FP32 element/contact/SDF kernels and FP32 K, C, M storage;
                                 rotations, residuals, norms, Krylov and Newmark stay FP64 */

typedef struct eb_model_desc {
    int32_t abi_version;          /* must be EB_ABI_VERSION */
    int32_t precision;            /* EB_PRECISION_* */
    int64_t n_nodes;
    int64_t n_beams;
    const double* X0;             /* 3*n_nodes  nodes.X₀ */
    const int64_t* conn;          /* 2*n_beams  (node1, node2) per beam, 1-based */
    const double* l0;             /* n_beams    beams.l₀ */
    const double* Re0;            /* 9*n_beams  beams.Rₑ⁰ (column-major) */
    const double* props;          /* EB_NPROPS*n_beams, beam-major */
    int64_t n_fixed;              /* bcs.fixed_dofs (1-based, strictly increasing) */
    const int64_t* fixed_dofs;
    int32_t n_gauss;              /* params.nᴳ_beams, 1..3 (ωᴳ_beams is a Vec3, src/def_params.jl:31-33) */
    int32_t pad0;
    const double* gauss_w;        /* params.ωᴳ_beams */
    const double* gauss_z;        /* params.zᴳ_beams */
    /* contact (RigidInteraction); surface = EB_SURFACE_NONE disables it */
    int32_t surface;
    int32_t sdf_prefiltered;      /* 1: sdf_samples already holds the padded quadratic B-spline
                                     coefficients, (n[0]+2)(n[1]+2)(n[2]+2) values x-fastest, as
                                     Interpolations.jl keeps them in parent(sitp.itp.coefs) */
    double kn, mu, eps_t, eta_n;  /* InteractionProperties kₙ, μ, εᵗ, ηₙ (kₜ, ηₜ, u̇ₛ are unused by the reference) */
    int64_t n_contact_beams;      /* BeamElementSurface.contact_beams (1-based) */
    const int64_t* contact_beams;
    double sphere_center[3];
    double sphere_radius;
    /* SDF: samples x-fastest (as read_vtk_sdf returns them), already sign-adjusted for `inside` */
    int64_t sdf_n[3];
    double sdf_d[3];              /* dx, dy, dz */
    double sdf_x0[3];             /* grid origin */
    const double* sdf_samples;    /* n[0]*n[1]*n[2] values; prefiltered inside the library */
} eb_model_desc;

/* Newton / time-integration parameters (SimulationParams, src/def_params.jl:1-35) plus the
 * device Krylov controls, which have no counterpart in the reference (it calls Pardiso). */
typedef struct eb_newton_params {
    double alpha, beta, gamma;    /* HHT α, β, γ */
    double tol_res, tol_disp;     /* tolerance_residual, tolerance_displacement */
    int32_t max_iterations;
    int32_t krylov_maxit;         /* BiCGStab iteration cap per linear solve (<= 0: default 200) */
    double krylov_tol;            /* BiCGStab relative residual target (<= 0: default 1e-11) */
} eb_newton_params;

/* state fields addressable by eb_get_field / eb_set_field */
enum eb_field {
    /* node state, current configuration (NodeBeam, src/beams/def_node_beam.jl:20-27) */
    EB_F_U = 0, EB_F_UDOT, EB_F_UDDOT, EB_F_W, EB_F_WDOT, EB_F_WDDOT, /* 3N each */
    EB_F_R, EB_F_DR,                                                  /* 9N each */
    /* node state at step n (def_node_beam.jl:30-37) */
    EB_F_U_N, EB_F_UDOT_N, EB_F_UDDOT_N, EB_F_W_N, EB_F_WDOT_N, EB_F_WDDOT_N,
    EB_F_R_N, EB_F_DR_N,
    /* Forces at n+1 and n (src/def_structures.jl:2-7), ndofs each */
    EB_F_FEXT, EB_F_TINT, EB_F_TK, EB_F_TC,
    EB_F_FEXT_N, EB_F_TINT_N, EB_F_TK_N, EB_F_TC_N,
    /* Solution (def_structures.jl:33-45), ndofs each */
    EB_F_D, EB_F_DDOT, EB_F_DDDOT, EB_F_RES, EB_F_DELTA_D,
    /* Matrices at n+1 and n, nzval in CSC order (nnz each) */
    EB_F_K, EB_F_C, EB_F_M, EB_F_K_N, EB_F_C_N, EB_F_M_N,
    EB_F_KTAN,                     /* Solution.Ktan nzval (nnz) */
    EB_F_RES_FREE,                 /* Solution.r_free (n_free) */
    EB_F_KTAN_FREE,                /* Solution.Ktan_free nzval (nnz_free) */
    EB_F_DELTA_D_FREE,             /* Solution.ΔD_free (n_free) */
    EB_F_IMPOSED,                  /* BoundaryConditions.imposed_displacements (n_fixed) */
    EB_F_ENERGY,                   /* Energy: strain, kinetic, contact (3) */
    EB_F_COUNT
};

typedef struct eb_ctx eb_ctx;

/* sizes of the assembled problem */
typedef struct eb_sizes {
    int64_t n_nodes, n_beams, ndofs, nnz, n_free, n_fixed, nnz_free;
} eb_sizes;

/* --- lifetime ----------------------------------------------------------------------------- */

/* Uploads the model, builds the CSC pattern, the element→nzval map and the free-dof gather
 * exactly as sparse_matrices_beams!/sparsity_beams (src/sparse_matrices.jl:6-104), prefilters
 * the SDF (src/interactions/def_interactions.jl:67-82) and zero-initialises every state field
 * except R, Rⁿ, ΔR, ΔRⁿ (identity), matching setup_state_simulation (src/inizialization.jl:17-32). */
int eb_create(const eb_model_desc* desc, eb_ctx** out);
int eb_destroy(eb_ctx* ctx);
const char* eb_last_error(const eb_ctx* ctx);
int eb_get_sizes(const eb_ctx* ctx, eb_sizes* out);

/* CSC pattern (colptr ndofs+1, rowval nnz, both 1-based), beams' global_sparsity_map
 * (144*n_beams, 1-based, column-major local index) and sparsity_free (nnz_free, 1-based);
 * any pointer may be NULL. Replaces sparsity_beams (src/sparse_matrices.jl:6-77). */
int eb_get_pattern(const eb_ctx* ctx, int64_t* colptr, int64_t* rowval, int64_t* sparsity_map,
                   int64_t* sparsity_free);

/* host <-> device copies of one state field; `count` must equal the field's length */
int eb_get_field(eb_ctx* ctx, int field, double* out, int64_t count);
int eb_set_field(eb_ctx* ctx, int field, const double* in, int64_t count);

/* --- oracle-comparable kernels (one reference function each) ------------------------------ */

/* assemble! (src/beams/assemble_beams.jl:1-70): zero K, C, M, Tⁱⁿᵗ, Tᵏ, Tᶜ and add every beam's
 * compute_beams (src/beams/compute_beams.jl:1-538) contribution; energies accumulate. */
int eb_assemble(eb_ctx* ctx);

/* assemble_contact! (src/interactions/compute&assemble_interactions_beams.jl:447-500):
 * Tᶜ += Tᶜₑ, K -= Kᶜₑ, C -= (1+α)Cᶜₑ over contact_beams. */
int eb_assemble_contact(eb_ctx* ctx, double alpha);

/* compute_tangent_and_residuals_corrector! / _predictor! (src/utils_solver.jl:80-89, :54-74) */
int eb_tangent_residuals_corrector(eb_ctx* ctx, double dt, double alpha, double beta, double gamma);
int eb_tangent_residuals_predictor(eb_ctx* ctx, double dt, double alpha, double beta, double gamma);

/* apply_boundary_conditions! (src/utils_bcs_and_loads.jl:5-29) */
int eb_apply_boundary_conditions(eb_ctx* ctx);

/* extract_free_dofs! (src/utils_solver.jl:94-102) */
int eb_extract_free_dofs(eb_ctx* ctx);

/* solve_free_dofs! (src/utils_solver.jl:107-120; Pardiso protocol src/linearsolvers.jl:2-28):
 * ΔD[free] = Ktan_free \ r_free. The device path is a chain-block-tridiagonal preconditioned
 * BiCGStab with FP64 dot products; `iterations` (may be NULL) receives its iteration count. */
int eb_solve_free_dofs(eb_ctx* ctx, const eb_newton_params* p, int* iterations);

/* initialize_global_corrector! / initialize_global_predictor! (src/utils_solver.jl:6-48) */
int eb_initialize_global_corrector(eb_ctx* ctx);
int eb_initialize_global_predictor(eb_ctx* ctx);

/* compute_global_corrector! / compute_global_predictor! (src/utils_solver.jl:178-190, :157-172) */
int eb_compute_global_corrector(eb_ctx* ctx, double dt, double beta, double gamma);
int eb_compute_global_predictor(eb_ctx* ctx, double dt, double beta, double gamma);

/* compute_local_corrector! / compute_local_predictor! (src/utils_solver.jl:224-245, :197-217) */
int eb_compute_local_corrector(eb_ctx* ctx, double dt, double beta, double gamma);
int eb_compute_local_predictor(eb_ctx* ctx, double dt, double beta, double gamma);

/* compute_norms_corrector (src/utils_solver.jl:126-151) */
int eb_compute_norms_corrector(eb_ctx* ctx, double* res_norm, double* dD_norm);

/* --- step-level entry points ---------------------------------------------------------------- */

/* update_loads! (src/utils_bcs_and_loads.jl:56-64): fᵉˣᵗ[dofs[i]] = values[i] (1-based dofs) */
int eb_update_loads(eb_ctx* ctx, int64_t n, const int64_t* dofs, const double* values);

/* update_boundary_conditions! (src/utils_bcs_and_loads.jl:42-50): imposed_displacements[:] = values */
int eb_update_boundary_conditions(eb_ctx* ctx, const double* values, int64_t n_fixed);

/* initialize_state_simulation! (src/inizialization.jl:37-45): assemble!, then matricesⁿ = matricesⁿ⁺¹ */
int eb_initialize_state_simulation(eb_ctx* ctx);

/* predictor! (src/predictor.jl:1-27) */
int eb_predictor(eb_ctx* ctx, const eb_newton_params* p, double dt);

/* corrector! (src/corrector.jl:1-59); *k receives the reference's return value (iterations + 1).
 * The whole Newton loop stays on the device; the host reads two norms per iteration. */
int eb_corrector(eb_ctx* ctx, const eb_newton_params* p, double dt, int* k);

/* solve_step_dynamics! (src/solve_step_dynamics.jl:1-14) without the load/BC update, which the
 * caller does through eb_update_loads / eb_update_boundary_conditions (they hold host closures) */
int eb_solve_step_dynamics(eb_ctx* ctx, const eb_newton_params* p, double dt, int* k);

/* update_converged! / update_not_converged! (src/utils_solver.jl:252-304) */
int eb_update_converged(eb_ctx* ctx);
int eb_update_not_converged(eb_ctx* ctx);

/* --- the time loop ---------------------------------------------------------------------------- */

/* Δt control and stop criterion of run_simulation! (SimulationParams, src/def_params.jl:1-35) */
typedef struct eb_time_params {
    double initial_timestep, min_timestep, max_timestep, simulation_end_time;
    int32_t accelerate_after_success_it;  /* Δt × 1.5 after this many successive steps with k <= 8 */
    int32_t stop_on_energy_threshold;     /* stop once the kinetic energy is below energy_threshold */
    double energy_threshold;
} eb_time_params;

/* Progress of the time loop, read and updated by eb_run_steps: before the first call t = 0,
 * dt = initial_timestep and everything else 0; later calls continue where the last one ended. */
typedef struct eb_run_state {
    double t, dt;                   /* tⁿ and the next step's Δt */
    int32_t successive;             /* successive converged steps since the last Δt change */
    int32_t stopped;                /* 1: the energy criterion ended the run; 2 (eb_run_steps_batch
                                       only): Δt fell below min_timestep, the instance aborted */
    int64_t steps, failed_steps;    /* converged steps, attempts repeated with Δt/2 */
    int64_t newton_iterations;      /* Σ (k - 1) over converged steps */
    int32_t last_k, pad;            /* k of the last attempt */
} eb_run_state;

/* Called before every step attempt with the attempt's time tⁿ⁺¹: sets the loads and imposed
 * displacements of that time (update_loads!, update_boundary_conditions!,
 * src/utils_bcs_and_loads.jl:42-64) through eb_update_loads / eb_update_boundary_conditions and
 * returns 0 (nonzero stops eb_run_steps with EB_ERR_ARG). NULL: no time-dependent loads. */
typedef int (*eb_step_callback)(eb_ctx* ctx, double t, void* user);

/* run_simulation!'s time loop (src/run_simulation.jl:38-112) inside the library, so that whole
 * time steps stay on the device: until max_steps more steps have converged, the end time is
 * reached or the energy criterion stops the run. Each attempt: callback(tⁿ + Δt), predictor!,
 * corrector!; k > max_iterations ⇒ Δt / 2 and update_not_converged! (EB_ERR_STATE when Δt falls
 * below min_timestep: the reference's "Simulation aborted."); otherwise Δt grows to
 * min(1.5 Δt, max_timestep) after accelerate_after_success_it successive steps with k <= 8, and
 * the step is committed (update_converged!). Wall-clock limits and output are the caller's (it
 * calls eb_run_steps in chunks). eb_initialize_state_simulation must have run. */
int eb_run_steps(eb_ctx* ctx, const eb_newton_params* p, const eb_time_params* tp, int64_t max_steps,
                 eb_step_callback callback, void* user, eb_run_state* state);

/* --- batches of independent structures (no counterpart in the reference) ------------------
 * For design sweeps: one context advances many structures (instances) together, each as if it
 * ran alone through run_simulation! (src/run_simulation.jl:1-119) with the batch's time step.
 * Every instance is a complete single-structure descriptor; precision, Gauss points, contact
 * surface kind and contact parameters must agree. Inside the context the instances are stored
 * one after the other: node, beam and dof numbers of instance s are shifted by the sizes of
 * instances 0..s-1, so every field, load dof and imposed-displacement list of the single-
 * structure API addresses the batch through those offsets. The corrector iterates every
 * instance until its own norms pass (corrector.jl:17-36); an instance that has converged is
 * frozen until the step ends, so its trajectory is the one it has alone. eb_solve_step_dynamics
 * returns the largest k of the instances (the step is repeated with Δt/2 for all when it exceeds
 * max_iterations). */
int eb_create_batch(const eb_model_desc* descs, int64_t n_instances, eb_ctx** out);
/* Called before every step attempt of instance `instance` with the attempt's time tⁿ⁺¹ (its own
 * time: instances of a batch advance with their own Δt); sets that instance's loads and imposed
 * displacements through eb_update_loads / eb_update_boundary_conditions (its dofs and fixed-dof
 * entries shifted by the sizes of the instances before it) and returns 0. NULL: no
 * time-dependent loads. */
typedef int (*eb_instance_callback)(eb_ctx* ctx, int64_t instance, double t, void* user);

/* run_simulation!'s time loop (src/run_simulation.jl:38-112) for every instance of a batch with
 * its own time and Δt: states[s] (n_instances records, as eb_run_steps keeps them) holds instance
 * s's progress. Each round, every instance that has not reached simulation_end_time and has not
 * stopped attempts one step, with its own Δt, until none of them has fewer than max_steps new
 * converged steps in this call (an instance that repeats a step does not hold the others up, so
 * some end with more than max_steps new steps); per instance, k > max_iterations halves Δt and repeats its step
 * (update_not_converged! for that instance alone), Δt below min_timestep ends that instance
 * (stopped = 2) without stopping the others, and a converged step is committed for that instance
 * (update_converged!) with the ×1.5 growth after accelerate_after_success_it easy steps.
 * Instances that do not attempt a step stay frozen. A linear solve that fails for one instance
 * (singular pivot block, breakdown, iteration cap) counts as that instance's Newton failure.
 * Each instance follows the trajectory it has alone under eb_run_steps. The energy stop is not
 * available (EB_ERR_ARG): the batch keeps one energy total. */
int eb_run_steps_batch(eb_ctx* ctx, const eb_newton_params* p, const eb_time_params* tp, int64_t max_steps,
                       eb_instance_callback callback, void* user, eb_run_state* states);

/* node_offsets and beam_offsets receive n_instances+1 entries (either may be NULL) */
int eb_get_instances(const eb_ctx* ctx, int64_t* n_instances, int64_t* node_offsets, int64_t* beam_offsets);
/* k of every instance at the end of the last corrector call (0: the instance took no part in
 * the step; max_iterations + 1: its Newton iteration or one of its linear solves failed) */
int eb_get_instance_iterations(const eb_ctx* ctx, int32_t* k_out);
/* relative residual and ‖ΔD‖ of every instance (2 per instance) from the last norms evaluation */
int eb_get_instance_norms(const eb_ctx* ctx, double* rn_dn);

/* --- instrumentation (no counterpart in the reference) ------------------------------------ */

/* cumulative counters since eb_create: [0] Newton iterations, [1] Krylov iterations,
 * [2] linear solves, [3] element-kernel launches */
int eb_get_counters(const eb_ctx* ctx, int64_t* out4);

/* per-kernel launch timing with HIP events recorded on the context's stream around every launch
 * of the kernels selected in `mask` (bit i = kernel id i); enabling resets the totals */
#define EB_KERNEL_BEAM 0          /* corotational element kernel */
#define EB_KERNEL_CONTACT 1       /* contact + SDF kernel */
#define EB_KERNEL_GATHER 2        /* assembly gathers (matrices, vectors, energies) */
#define EB_KERNEL_PRECOND 3       /* chain factorisation */
#define EB_KERNEL_KRYLOV 4        /* one BiCGStab iteration (all its kernels) */
#define EB_KERNEL_UPDATE 5        /* Newmark/rotation update and norms */
#define EB_KERNEL_GATHER_MATRICES 6   /* the tangent's gather alone (part of EB_KERNEL_GATHER) */
#define EB_KERNEL_COUNT 7
int eb_profile_enable(eb_ctx* ctx, uint32_t mask);
/* total milliseconds and number of timed launches of kernel `id` since enabling */
int eb_profile_read(eb_ctx* ctx, int id, double* total_ms, int64_t* launches);

/* beams listed in contact by the detection and their Gauss points in contact, summed over the
 * contact-tangent launches since eb_create (out2[0], out2[1]): the work of those launches */
int eb_get_contact_stats(eb_ctx* ctx, int64_t* out2);
/* the work of the launches timed since eb_profile_enable (out3): [0] beams through the element
 * kernel, [1] tangent entries through the matrix gather, [2] chain positions factored, each summed
 * over the instances taking part in each launch (instances frozen in the Newton iteration are
 * skipped by the kernels, so a launch's work is that of its active instances). Counted only for
 * the kernel groups eb_profile_enable times (EB_KERNEL_BEAM, EB_KERNEL_GATHER_MATRICES,
 * EB_KERNEL_PRECOND); eb_profile_enable resets them. */
int eb_get_work_counters(eb_ctx* ctx, int64_t* out3);

/* --- output ----------------------------------------------------------------------------------- */

/* recompute_at_gausspts_beams! (src/IO/visualization_beams.jl:132-181), the data of write_VTK:
 * per beam, n_points points evenly spaced along it (range(-1, 1, length = n_points); the
 * reference uses 10), from the current node state: position N₁x₁ + N₂x₂ + Rₑuᵗ (with the cubic
 * bending offset uᵗ), displacement N₁u₁ + N₂u₂, velocity N₁u̇₁ + N₂u̇₂, strain 1 - lₙ/l₀ and stress
 * E(1 - lₙ/l₀). Point i of beam b (both 0-based) is entry b·n_points + i; vectors take 3 doubles
 * per point. Any output pointer may be NULL. */
int eb_output_points(eb_ctx* ctx, int32_t n_points, double* position, double* displacement, double* velocity,
                     double* strain, double* stress);

/* eb_output_points without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight output loops, harnesses that
 * forbid hipMalloc/hipFree): the same points and fields, staged in a fixed device array of
 * EB_NOALLOC_SCRATCH_DOUBLES doubles that every GPU the library runs on holds until the process
 * ends. The requested outputs may take at most that many doubles together (n_points × n_beams ×
 * 3 per requested vector field, × 1 per requested scalar field); larger requests fail with
 * EB_ERR_ARG. Calls on contexts of one GPU share that GPU's array: calls that overlap in time on
 * one GPU (several host threads) race; calls in sequence do not. */
#define EB_NOALLOC_SCRATCH_DOUBLES (1 << 22)
int eb_output_points_noalloc(eb_ctx* ctx, int32_t n_points, double* position, double* displacement, double* velocity,
                             double* strain, double* stress);

/* Prefiltered quadratic B-spline coefficients of instance `instance`'s signed distance field as
 * the device holds them ((n0+2)(n1+2)(n2+2), x fastest, converted to double) — the coefficient
 * array of the reference's interpolant (src/interactions/def_interactions.jl:77). The device
 * computes them at eb_create from the samples (unless the descriptor passes them prefiltered). */
int eb_get_sdf_coefficients(eb_ctx* ctx, int64_t instance, double* out, int64_t count);

/* --- execution mode and launch shapes (no counterpart in the reference) --------------------- */

/* How eb_corrector issues a Newton iteration. EAGER: the host launches every kernel.
 * GRAPH: the iteration (assembly, lift, factorisation, one BiCGStab iteration, the update and
 * norms gated on the solve's done flag, the two status copies) is captured once as a HIP graph
 * per set of step parameters (Δt, α, β, γ, tolerances) and replayed; results are identical.
 * FUSED: eb_solve_step_dynamics (and the time loops built on it) runs the predictor as one launch
 * and each Newton iteration of the corrector as two: the element kernel, then one workgroup going
 * through the gathers, lift, solve, updates and norms on the device code of the EAGER path, so the
 * results are the EAGER path's, bit for bit. The iterations of a step are launched in batches (the
 * previous step's count, then doubling), the host waiting once per batch; an iteration launched
 * after convergence returns at once. For a context of one structure of at most 4096 beams whose chain preconditioner holds every
 * coupling (open chains, rings), with the block-LU chain algorithm and no learned start;
 * eb_set_exec_mode refuses it for any other context (EB_ERR_ARG, reason in eb_last_error).
 * eb_predictor and eb_corrector called on their own keep the EAGER launches. */
#define EB_EXEC_EAGER 0
#define EB_EXEC_GRAPH 1
#define EB_EXEC_FUSED 2
int eb_set_exec_mode(eb_ctx* ctx, int mode);
/* mode, graph captures and graph replays since eb_create (any pointer may be NULL) */
int eb_get_exec_stats(const eb_ctx* ctx, int32_t* mode, int64_t* captures, int64_t* replays);

/* Launch shapes chosen by eb_autotune (0 in eb_set_tuning: the built-in default):
 *   EB_TUNE_CHAIN_LDS_KB    LDS budget (KB) of the reduced chain in the chain factorisation
 *                           (candidates 16, 32, 48, 96; default 48; at most 96, and at most what
 *                           the device grants a workgroup: a larger budget fails the factorisation)
 *   EB_TUNE_CHAIN_SOLVE_LDS chain solve work vectors in LDS (1, default) or global memory (2)
 *   EB_TUNE_GATHER_BLOCK    workgroup size of the matrix gather (64, 128, 256; default 256)
 *   EB_TUNE_CHAIN_ALGO      how the chain preconditioner factors and solves each chain:
 *                           EB_CHAIN_CYCLIC_REDUCTION (one workgroup per chain, ceil(log2 n)
 *                           parallel levels) or EB_CHAIN_BLOCK_LU (default: six lanes per chain,
 *                           n sequential block steps, ten chains per wavefront)
 * The first three only change launch shapes: results are bit-identical whatever they are. The
 * chain algorithm changes the preconditioner's rounding: results agree to the Krylov tolerance
 * and the Newton iteration counts are the same, but they are not bit-identical. */
#define EB_TUNE_CHAIN_LDS_KB 0
#define EB_TUNE_CHAIN_SOLVE_LDS 1
#define EB_TUNE_GATHER_BLOCK 2
#define EB_TUNE_CHAIN_ALGO 3
#define EB_TUNE_COUNT 4
#define EB_TUNE_CANDIDATES 4
#define EB_CHAIN_CYCLIC_REDUCTION 1
#define EB_CHAIN_BLOCK_LU 2
/* Times every candidate on this context's sizes and current tangent (call after a time step)
 * and keeps the fastest of each; choice[EB_TUNE_COUNT] receives them and, when not NULL,
 * ms_per_candidate[EB_TUNE_COUNT * EB_TUNE_CANDIDATES] the average launch time of each
 * candidate (-1: not a candidate; the chain algorithm is timed as one factorisation plus one
 * application). */
int eb_autotune(eb_ctx* ctx, int32_t* choice, double* ms_per_candidate);
int eb_set_tuning(eb_ctx* ctx, const int32_t* choice);
int eb_get_tuning(const eb_ctx* ctx, int32_t* choice);

#ifdef __cplusplus
}
#endif

#endif /* ENDOBEAMS_HIP_H */
