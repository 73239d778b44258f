// eb_fused.h — the predictor (src/predictor.jl:1-27) and the part of each corrector iteration
// (src/corrector.jl:15-53) after the element and contact kernels — the gathers into the tangent and
// the HHT residual, the energies, the boundary-condition lift, the linear solve, the Newmark and
// rotation updates and the norms — each as one launch of one 256-thread workgroup, phase after phase
// with workgroup barriers in between, for a small structure. Every phase calls the device body of
// the kernel the multi-kernel path launches for it, with the block indices that launch would have,
// so each result is the multi-kernel path's. Such a structure otherwise spends a Newton iteration on
// some twenty dependent launches of a few microseconds each; here an iteration is the element
// kernel (and the contact kernels) and one launch. The element code stays in its own kernels
// (compiled with structural-zero folding, see eb_element.hip): inside these kernels its registers
// would be live beside the solver's.
#pragma once
#include "eb_beam_kernel.h"
#include "eb_linalg.h"

namespace ebd {

constexpr int FUSED_THREADS = 256;
constexpr int64_t FUSED_MAX_BEAMS = 4096;   // one workgroup gathers, solves and updates the structure

template <typename S> struct FusedArgs {
    // the element staging (written by the element and contact kernels)
    BeamStage<S> st;
    const double* sTc;
    int64_t E, N, ndofs, nnz;
    // assembly and the HHT combination
    const int32_t* blk_ptr;
    const uint32_t* blk_src;
    const int32_t* inc_ptr;
    const uint32_t* inc_src;
    const double *Kn, *Cn, *Mn;
    double* Ktan;
    const StepCoef* coef;
    double *Tint, *Tk, *Tc, *fext, *fext_n, *Tc_n, *Tint_n, *Tk_n, *res, *energy, *partials;
    int energy_grid;
    // boundary conditions and the Newmark / rotation updates
    const int32_t *bptr, *bcol, *fixidx;
    const double* imposed;
    const uint8_t* is_fixed;
    int64_t n_fixed;
    NodeFields nf;
    double *D, *Dd, *Ddd, *dD, *pdD;
    // the linear solve: chain block LU (with the closures of rings), BiCGStab
    ChainDev cd;
    ClosureDev cl;
    int32_t n_closures;
    int64_t n_chains;
    Segs sg;
    KrylovCtl* kc;
    KrylovAll* kall;
    int32_t* pivot_inst;
    const uint8_t* inst_exact;
    double *kx, *kr, *krhat, *kp, *kv, *ky, *ks, *kz, *kt;
    double krylov_tol;
    int krylov_maxit;
    // Newton control
    Active act;
    uint8_t* active;    // act.flag, writable
    const uint8_t* part;
    uint8_t* pfail;
    int32_t* kiter;
    int32_t* kdone;
    NewtonAll* nall;
    double* norms;
    double alpha;
    NewtonCtl nc;
    int reuse_factor;   // the predictor's tangent and factors are those of the last step (predictor_reuses_factor)
    int64_t* out;       // [3]: the predictor's Krylov iterations; corrector iterations run; their Krylov iterations
};

// a flag another wavefront of the workgroup wrote, read past the caches a uniform load may use
EBD int fused_flag(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

// blocks of an element-wise body (256 items each, as its launch)
template <class F> EBD void fused_each(int64_t n, F&& f) {
    const int64_t nb = (n + FUSED_THREADS - 1) / FUSED_THREADS;
    for (int64_t j = 0; j < nb; ++j) f(VB{j, nb, 0});
    __syncthreads();
}
// blocks of a body that reduces through LDS: a barrier after each block
template <class F> EBD void fused_blocks(int64_t nb, F&& f) {
    for (int64_t j = 0; j < nb; ++j) {
        f(VB{j, nb, 0});
        __syncthreads();
    }
}
// blocks of a 64-lane body (the chain block-LU kernels): wavefront 0 runs them, in block order
template <class F> EBD void fused_wave0(int64_t nb, F&& f) {
    if (threadIdx.x < 64)
        for (int64_t j = 0; j < nb; ++j) f(VB{j, nb, 0});
    __syncthreads();
}
EBD int64_t lu_blocks(int64_t items) { return (items + LU_CHAINS - 1) / LU_CHAINS; }
constexpr VB ONE_BLOCK{0, 1, 0};

// precond_apply of the multi-kernel path (block LU, no crossing cover): x = M⁻¹ b
template <typename S> EBD void fused_precond(const FusedArgs<S>& a, const double* b, double* x, int direct_only) {
    fused_wave0(lu_blocks(a.n_chains), [&](const VB& vb) { chain_lu_solve_b<false>(a.cd, b, x, a.kc, direct_only, vb); });
    fused_blocks(a.n_closures, [&](const VB& vb) { closure_apply_b(a.cd, a.cl, x, a.kc, direct_only, vb); });
}

// solve_start + first_iteration (the direct step) + solve_finish of the multi-kernel path;
// returns the BiCGStab iterations launched (the direct step counts as one)
template <typename S> EBD int fused_solve(const FusedArgs<S>& a, bool factor) {
    const Segs& sg = a.sg;
    const int64_t nc = sg.nchunks;
    if (factor) {
        solve_begin_b((int)sg.ninst, a.pivot_inst, a.kall, ONE_BLOCK);
        __syncthreads();
        fused_wave0(lu_blocks(a.n_chains), [&](const VB& vb) { chain_lu_factor_b(a.cd, a.Ktan, a.pivot_inst, a.act.flag, vb); });
        if (a.n_closures > 0) {
            fused_wave0(lu_blocks(12 * (int64_t)a.n_closures), [&](const VB& vb) { closure_columns_lu_b(a.cd, a.cl, a.act.flag, vb); });
            fused_blocks(a.n_closures, [&](const VB& vb) { closure_setup_b(a.cd, a.cl, a.Ktan, a.act.flag, vb); });
        }
    }
    fused_blocks(nc, [&](const VB& vb) {
        krylov_init_b(sg, a.is_fixed, a.res, a.kx, a.kr, a.krhat, a.kp, a.kv, a.partials, vb);
    });
    krylov_init_fin_b(a.partials, sg, a.kc, a.krylov_tol, a.act.flag, a.pivot_inst, a.inst_exact, a.kall, ONE_BLOCK);
    __syncthreads();
    // the direct step x = M⁻¹ b and its residual
    fused_precond(a, a.krhat, a.kx, 1);
    fused_blocks(nc, [&](const VB& vb) {
        krylov_direct_residual_b(sg, a.bptr, a.bcol, a.is_fixed, a.Ktan, a.kx, a.krhat, a.kr, a.partials, a.kc, vb);
    });
    krylov_direct_fin_b(a.partials, sg, a.kc, a.kall, ONE_BLOCK);
    __syncthreads();
    int launched = 1;
    while (!fused_flag(&a.kall->all_done) && launched < a.krylov_maxit) {
        fused_blocks(nc, [&](const VB& vb) { krylov_p_b(sg, a.kc, a.kr, a.kp, a.kv, vb); });
        fused_precond(a, a.kp, a.ky, 0);
        fused_blocks(nc, [&](const VB& vb) {
            spmv_masked_b<1>(sg, a.bptr, a.bcol, a.is_fixed, a.Ktan, a.ky, a.kv, a.krhat, a.partials, a.kc, vb);
        });
        krylov_alpha_fin_b(a.partials, sg, a.kc, ONE_BLOCK);
        __syncthreads();
        fused_blocks(nc, [&](const VB& vb) { krylov_s_b(sg, a.kc, a.kr, a.kv, a.ky, a.ks, a.kx, a.partials, vb); });
        krylov_s_fin_b(a.partials, sg, a.kc, ONE_BLOCK);
        __syncthreads();
        fused_precond(a, a.ks, a.kz, 0);
        fused_blocks(nc, [&](const VB& vb) {
            spmv_masked_b<2>(sg, a.bptr, a.bcol, a.is_fixed, a.Ktan, a.kz, a.kt, a.ks, a.partials, a.kc, vb);
        });
        krylov_omega_fin_b(a.partials, sg, a.kc, ONE_BLOCK);
        __syncthreads();
        fused_blocks(nc, [&](const VB& vb) {
            krylov_x_b(sg, a.kc, a.kz, a.ks, a.kt, a.krhat, a.kx, a.kr, a.partials, vb);
        });
        krylov_rho_fin_b(a.partials, sg, a.kc, a.kall, ONE_BLOCK);
        __syncthreads();
        ++launched;
    }
    return launched;
}

// bc_lift of the multi-kernel path
template <typename S> EBD void fused_lift(const FusedArgs<S>& a) {
    if (a.n_fixed == 0) return;
    fused_each(a.ndofs, [&](const VB& vb) {
        bc_lift_b(a.ndofs, a.bptr, a.bcol, a.fixidx, a.imposed, a.D, a.Ktan, a.dD, a.res, a.act, vb);
    });
}

// the gathers of assemble_all (tangent and corrector residual) and the energies, on the staging the
// element (and contact) kernels just wrote
template <typename S, bool CONTACT> EBD void fused_gathers(const FusedArgs<S>& a) {
    const double c1 = 1 + a.alpha;
    fused_each(a.nnz / 4, [&](const VB& vb) {
        gather_matrices_b<S>(a.nnz, a.blk_ptr, a.blk_src, a.st.K, a.st.C, a.st.M, nullptr, nullptr, nullptr, a.Ktan, c1, a.coef,
                             1, 0, a.act, vb);
    });
    fused_each(a.ndofs, [&](const VB& vb) {
        gather_vectors_b<S>(a.ndofs, a.inc_ptr, a.inc_src, a.st.Tint, a.st.Tk, a.sTc, CONTACT ? 1 : 0, a.Tint, a.Tk, a.Tc, 1,
                            a.alpha, a.fext, a.fext_n, a.Tc_n, a.Tint_n, a.res, a.act, vb);
    });
    fused_blocks(a.energy_grid, [&](const VB& vb) { energy_partials_b(a.E, a.st.energy, a.partials, CONTACT ? 3 : 1, a.act, vb); });
    energy_finalize_b(a.partials, a.energy_grid, a.energy, ONE_BLOCK);
    __syncthreads();
}

// predictor! of one structure (predictor_impl), then the start of its corrector loop (init_global
// and newton_reset of corrector_impl)
template <typename S> __global__ void __launch_bounds__(FUSED_THREADS) fused_predictor(FusedArgs<S> a) {
    const int ninst = (int)a.sg.ninst;
    const int32_t kfail = a.nc.max_iterations + 1;
    for (int s = threadIdx.x; s < ninst; s += FUSED_THREADS) a.pfail[s] = 0;
    __syncthreads();
    newton_reset_b(ninst, a.part, nullptr, a.active, a.kiter, a.kdone, 0, ONE_BLOCK);
    __syncthreads();
    fused_each(a.N, [&](const VB& vb) { init_global_b(a.N, a.nf, a.D, a.Dd, a.Ddd, a.dD, 1, a.act, vb); });
    if (!a.reuse_factor)
        fused_each(a.nnz, [&](const VB& vb) {
            combine_tangent_b(a.nnz, a.Kn, a.Cn, a.Mn, a.Ktan, 1 + a.alpha, a.coef, a.act.blk_inst, vb);
        });
    fused_each(a.ndofs, [&](const VB& vb) {
        residual_predictor_b(a.ndofs, a.bptr, a.bcol, a.alpha, a.coef, a.act.node_inst, a.fext, a.Tint_n, a.Tk_n, a.fext_n,
                             a.Dd, a.Ddd, a.Cn, a.Mn, a.res, vb);
    });
    fused_lift(a);
    fused_solve(a, !a.reuse_factor);
    if (threadIdx.x == 0) a.out[0] = fused_flag(&a.kall->max_iters);
    solve_screen_b(ninst, a.kc, a.active, a.kiter, 0, a.pfail, nullptr, ONE_BLOCK);
    __syncthreads();
    fused_each(a.ndofs, [&](const VB& vb) { scatter_solution_b(a.ndofs, a.is_fixed, a.kx, a.dD, a.act, nullptr, vb); });
    fused_each(a.ndofs, [&](const VB& vb) { global_predictor_b(a.ndofs, a.coef, a.D, a.Dd, a.Ddd, a.dD, a.act, vb); });
    fused_each(a.N, [&](const VB& vb) { local_predictor_b(a.N, a.nf, a.D, a.Dd, a.Ddd, a.coef, a.act, vb); });
    for (int64_t i = threadIdx.x; i < a.ndofs; i += FUSED_THREADS) a.pdD[i] = a.dD[i];
    __syncthreads();
    // corrector_impl's start: the corrector's global state, the instances taking part active (one whose
    // predictor solve failed frozen with k = max_iterations + 1)
    fused_each(a.N, [&](const VB& vb) { init_global_b(a.N, a.nf, a.D, a.Dd, a.Ddd, a.dD, 0, a.act, vb); });
    newton_reset_b(ninst, a.part, a.pfail, a.active, a.kiter, a.kdone, kfail, ONE_BLOCK);
    // the corrector loop runs while instances take part (corrector_impl counts them on the host)
    if (threadIdx.x == 0) {
        int n = 0;
        for (int s = 0; s < ninst; ++s) n += a.part[s] ? 1 : 0;
        a.nall->n_active = n;
        a.out[1] = 0;
        a.out[2] = 0;
    }
}

// one corrector iteration after the element and contact kernels (newton_body with the solve carried
// to its end): the gathers, lift, solve, screen, ΔD, updates and norms with the convergence bookkeeping
// The host launches iterations ahead without waiting: one that finds no instance active (the loop
// has ended) returns at once, so the iterations that run are those of corrector_impl's loop.
template <typename S, bool CONTACT> __global__ void __launch_bounds__(FUSED_THREADS) fused_iteration(FusedArgs<S> a) {
    if (fused_flag(&a.nall->n_active) == 0) return;   // uniform over the workgroup
    const int ninst = (int)a.sg.ninst;
    const int32_t kfail = a.nc.max_iterations + 1;
    fused_gathers<S, CONTACT>(a);
    fused_lift(a);
    fused_solve(a, true);
    solve_screen_b(ninst, a.kc, a.active, a.kiter, kfail, nullptr, nullptr, ONE_BLOCK);
    __syncthreads();
    fused_each(a.ndofs, [&](const VB& vb) { scatter_solution_b(a.ndofs, a.is_fixed, a.kx, a.dD, a.act, nullptr, vb); });
    fused_each(a.ndofs, [&](const VB& vb) { global_corrector_b(a.ndofs, a.coef, a.D, a.Dd, a.Ddd, a.dD, a.act, nullptr, vb); });
    fused_each(a.N, [&](const VB& vb) { local_corrector_b(a.N, a.nf, a.D, a.Dd, a.Ddd, a.dD, a.coef, a.act, nullptr, vb); });
    fused_blocks(a.sg.nchunks, [&](const VB& vb) {
        norms_partials_b(a.sg, a.is_fixed, a.dD, a.res, a.fext, a.Tc, a.Tk, a.partials, vb);
    });
    norms_finalize_b(a.partials, a.sg, a.norms, a.nc, a.active, a.kiter, a.nall, nullptr, a.kdone, ONE_BLOCK);
    if (threadIdx.x == 0) {
        a.out[1] += 1;
        a.out[2] += fused_flag(&a.kall->max_iters);
    }
}

}  // namespace ebd
