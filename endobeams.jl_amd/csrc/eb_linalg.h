// eb_linalg.h — assembly gathers, HHT-α combine, boundary-condition lift, masked block-row SpMV,
// chain block-tridiagonal preconditioner, BiCGStab kernels, Newmark/rotation updates and FP64
// reductions for the device Newton iteration.
#pragma once
#include "eb_dmath.h"

namespace ebd {

constexpr int RED_BLOCK = 256;
constexpr int RED_MAX_GRID = 1024;

// six doubles at a 16-byte aligned address, in three 16-byte accesses
EBD void ld6(const double* p, double* out) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double2 v = q[k];
        out[2 * k] = v.x;
        out[2 * k + 1] = v.y;
    }
}
EBD void st6(double* p, const double* v) {
    double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// Every kernel below is a device body taking the index of the block it works on (VB: blockIdx.x,
// gridDim.x, blockIdx.y) and a __global__ wrapper passing the launch's own; a kernel that runs
// several of these phases in one workgroup calls the bodies with the block indices it iterates.
struct VB { int64_t x, n, y; };
#define EB_VB (VB{(int64_t)blockIdx.x, (int64_t)gridDim.x, (int64_t)blockIdx.y})
// EB_BODIES_ONLY: the bodies without their __global__ wrappers (a translation unit that runs the
// bodies from a kernel of its own, beside the one that launches the wrappers)

// ---- deterministic reductions ---------------------------------------------------------------

template <int NQ> EBD void block_sum_store(double (&v)[NQ], double* partials, int grid, int64_t bid) {
    __shared__ double sh[NQ][RED_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0;
            for (int w = 0; w < RED_BLOCK / 64; ++w) s += sh[q][w];
            partials[q * RED_MAX_GRID + bid] = s;
        }
    }
}

// sums partials[q][0..grid) in fixed order inside one 256-thread block; result to out[q]
template <int NQ> EBD void finalize_sum(const double* partials, int grid, double (&out)[NQ]) {
    __shared__ double sh[NQ][RED_BLOCK / 64];
    __shared__ double res[NQ];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = 0;
        for (int i = threadIdx.x; i < grid; i += RED_BLOCK) x += partials[q * RED_MAX_GRID + i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0;
            for (int w = 0; w < RED_BLOCK / 64; ++w) s += sh[q][w];
            res[q] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = res[q];
}

// ---- per-instance (segmented) reductions ---------------------------------------------------
// A context holds one or more independent structures (instances) laid out one after the other.
// Reductions over dofs run over chunks that never straddle two instances: chunk c covers dofs
// [lo[c], hi[c]) of instance inst[c], and instance s owns chunks cptr[s] .. cptr[s+1]-1. One
// workgroup reduces one chunk; a finalize kernel sums each instance's chunk partials in chunk
// order, so every sum is deterministic.
constexpr int SEG_CHUNK = 1024;   // dofs per chunk
constexpr int FIN_BLOCK = 1024;   // threads of the per-instance finalize kernels

struct Segs {
    const int64_t *lo, *hi;
    const int32_t* inst;
    const int32_t* cptr;
    int32_t nchunks, ninst;
};

template <int NQ> EBD void seg_block_store(double (&v)[NQ], double* partials, int nchunks, int64_t bid) {
    __shared__ double sh[NQ][RED_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double t = 0;
            for (int w = 0; w < RED_BLOCK / 64; ++w) t += sh[q][w];
            partials[(int64_t)q * nchunks + bid] = t;
        }
    }
}

template <int NQ> EBD void seg_sum(const double* partials, const Segs& sg, int s, double (&out)[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = 0;
    for (int c = sg.cptr[s]; c < sg.cptr[s + 1]; ++c) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[q] += partials[(int64_t)q * sg.nchunks + c];
    }
}

// Newton activity per instance: an instance that has converged is frozen (no assembly, solve or
// update) until the step ends, so that it ends exactly where it would end alone.
struct Active {
    const uint8_t* flag;       // [ninst]
    const int32_t* node_inst;  // [N]
    const int32_t* beam_inst;  // [E]
    const int32_t* blk_inst;   // [nnz/36]
};

// Per-instance coefficients of the HHT tangent and the Newmark/rotation updates with the
// instance's own Δt (β, γ shared by the batch), computed on the host with the expressions of
// utils_solver.jl:54-245 and read through the instance tables, so that instances of one batch may
// advance with different time steps.
struct HHTc { double c_w, c_wd, c_wdd, d_w, d_wd, d_wdd; };  // ẇ = ΔR(c_w w + c_wd ẇⁿ + c_wdd ẅⁿ), ẅ likewise
struct StepCoef {
    double dt;
    double gbdt;   // γ/(βΔt): the tangent's C factor, Ḋ update
    double inv;    // 1/(βΔt²): the tangent's M factor, D̈ update
    double gob;    // γ/β: predictor velocity terms
    double dt2b;   // Δt(2β - γ)/(2β): global predictor
    double t2c;    // ((Δt/2)(2β - γ))/β: predictor residual
    HHTc h;        // local (rotation) updates
};

// ---- assembly gathers (atomic-free, beam order = the reference's serial scatter order) ------

// four consecutive staging values (16-byte aligned: a block's 36 entries start on a multiple of
// four) widened to double
EBD void ld4w(const float* __restrict__ p, double (&o)[4]) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    o[0] = double(v.x); o[1] = double(v.y); o[2] = double(v.z); o[3] = double(v.w);
}
EBD void ld4w(const double* __restrict__ p, double (&o)[4]) {
    const double2 v0 = *reinterpret_cast<const double2*>(p), v1 = *reinterpret_cast<const double2*>(p + 2);
    o[0] = v0.x; o[1] = v0.y; o[2] = v1.x; o[3] = v1.y;
}
EBD void st4(double* p, const double (&v)[4]) {
    *reinterpret_cast<double2*>(p) = double2{v[0], v[1]};
    *reinterpret_cast<double2*>(p + 2) = double2{v[2], v[3]};
}

// K, C, M and/or Ktan = (1+α)K + γ/(βΔt) C + M/(βΔt²) for every block-row entry; one lane per
// four consecutive entries of a 6×6 block (nnz/4 lanes), each entry summed over the block's
// contributions in beam order
template <typename P>
EBD void gather_matrices_b(int64_t nnz, const int32_t* __restrict__ blk_ptr,
                                                       const uint32_t* __restrict__ blk_src, const P* __restrict__ sK,
                                                       const P* __restrict__ sC, const P* __restrict__ sM, double* K,
                                                       double* C, double* M, double* Ktan, double c1,
                                                       const StepCoef* __restrict__ coef, int with_tangent, int write_kcm,
                                                       Active act, const VB& vb) {
    const int64_t q = 4 * ((int64_t)vb.x * blockDim.x + threadIdx.x);
    if (q >= nnz) return;
    const int64_t b = q / 36;
    if (!act.flag[act.blk_inst[b]]) return;
    const int r = (int)(q % 36);   // row-major position 6a + c in the block, a multiple of four
    double k[4] = {0, 0, 0, 0}, cc[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
    for (int s = blk_ptr[b]; s < blk_ptr[b + 1]; ++s) {
        const uint32_t src = blk_src[s];
        const int64_t e = src >> 2;
        const int sr = (src >> 1) & 1, sc = src & 1;
        const int64_t idx = e * 144 + (2 * sr + sc) * 36 + r;   // stage_idx(e, 6 sr + a, 6 sc + c)
        double vk[4], vc[4], vm[4];
        ld4w(sK + idx, vk);
        ld4w(sC + idx, vc);
        ld4w(sM + idx, vm);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            k[j] += vk[j];
            cc[j] += vc[j];
            m[j] += vm[j];
        }
    }
    if (write_kcm) {
        st4(K + q, k);
        st4(C + q, cc);
        st4(M + q, m);
    }
    if (with_tangent) {
        const StepCoef& sc = coef[act.blk_inst[b]];
        const double c2 = sc.gbdt, c3 = sc.inv;
        double t[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] = (c1 * k[j] + c2 * cc[j]) + c3 * m[j];
        st4(Ktan + q, t);
    }
}
#ifndef EB_BODIES_ONLY
template <typename P>
__global__ void __launch_bounds__(256) gather_matrices(int64_t nnz, const int32_t* __restrict__ blk_ptr,
                                                       const uint32_t* __restrict__ blk_src, const P* __restrict__ sK,
                                                       const P* __restrict__ sC, const P* __restrict__ sM, double* K,
                                                       double* C, double* M, double* Ktan, double c1,
                                                       const StepCoef* __restrict__ coef, int with_tangent, int write_kcm,
                                                       Active act) { gather_matrices_b<P>(nnz, blk_ptr, blk_src, sK, sC, sM, K, C, M, Ktan, c1, coef, with_tangent, write_kcm, act, EB_VB); }
#endif

// Tⁱⁿᵗ, Tᵏ, Tᶜ per dof, optionally followed by the corrector residual (utils_solver.jl:87)
template <typename P>
EBD void gather_vectors_b(int64_t ndofs, const int32_t* __restrict__ inc_ptr,
                                                      const uint32_t* __restrict__ inc_src, const double* __restrict__ sTint,
                                                      const double* __restrict__ sTk, const double* __restrict__ sTc, int with_contact,
                                                      double* Tint, double* Tk, double* Tc, int with_residual, double alpha,
                                                      const double* fext, const double* fext_n, const double* Tc_n,
                                                      const double* Tint_n, double* res, Active act, const VB& vb) {
    int64_t d = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (d >= ndofs) return;
    const int64_t n = d / 6;
    if (!act.flag[act.node_inst[n]]) return;
    const int c = (int)(d % 6);
    double ti = 0, tk = 0, tc = 0;
    for (int s = inc_ptr[n]; s < inc_ptr[n + 1]; ++s) {
        uint32_t src = inc_src[s];
        int64_t e = src >> 1;
        int slot = src & 1;
        int64_t idx = e * 12 + 6 * slot + c;
        tk += sTk[idx];
        ti += sTint[idx];
        if (with_contact) tc += sTc[idx];
    }
    Tint[d] = ti;
    Tk[d] = tk;
    Tc[d] = tc;
    if (with_residual)
        res[d] = ((1 + alpha) * ((fext[d] + tc) - ti) - alpha * ((fext_n[d] + Tc_n[d]) - Tint_n[d])) - tk;
}
#ifndef EB_BODIES_ONLY
template <typename P>
__global__ void __launch_bounds__(256) gather_vectors(int64_t ndofs, const int32_t* __restrict__ inc_ptr,
                                                      const uint32_t* __restrict__ inc_src, const double* __restrict__ sTint,
                                                      const double* __restrict__ sTk, const double* __restrict__ sTc, int with_contact,
                                                      double* Tint, double* Tk, double* Tc, int with_residual, double alpha,
                                                      const double* fext, const double* fext_n, const double* Tc_n,
                                                      const double* Tint_n, double* res, Active act) { gather_vectors_b<P>(ndofs, inc_ptr, inc_src, sTint, sTk, sTc, with_contact, Tint, Tk, Tc, with_residual, alpha, fext, fext_n, Tc_n, Tint_n, res, act, EB_VB); }
#endif

// energies: Σ over beams of [E][3] staging, added to the running totals (never reset, as in
// the reference's Energy accumulation, assemble_beams.jl:56-57)
EBD void energy_partials_b(int64_t E, const double* st, double* partials, int mask,
                                                             Active act, const VB& vb) {
    double v[3] = {0, 0, 0};
    for (int64_t e = (int64_t)vb.x * RED_BLOCK + threadIdx.x; e < E; e += (int64_t)vb.n * RED_BLOCK) {
        if (!act.flag[act.beam_inst[e]]) continue;
        if (mask & 1) { v[0] += st[3 * e]; v[1] += st[3 * e + 1]; }
        if (mask & 2) v[2] += st[3 * e + 2];
    }
    block_sum_store<3>(v, partials, vb.n, vb.x);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) energy_partials(int64_t E, const double* st, double* partials, int mask,
                                                             Active act) { energy_partials_b(E, st, partials, mask, act, EB_VB); }
#endif
EBD void energy_finalize_b(const double* partials, int grid, double* energy, const VB& vb) {
    double out[3];
    finalize_sum<3>(partials, grid, out);
    if (threadIdx.x == 0) { energy[0] += out[0]; energy[1] += out[1]; energy[2] += out[2]; }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) energy_finalize(const double* partials, int grid, double* energy) { energy_finalize_b(partials, grid, energy, EB_VB); }
#endif

// ---- HHT combine and residuals (src/utils_solver.jl:54-89) ----------------------------------

EBD void combine_tangent_b(int64_t nnz, const double* K, const double* C, const double* M,
                                                       double* Ktan, double c1, const StepCoef* __restrict__ coef,
                                                       const int32_t* __restrict__ blk_inst, const VB& vb) {
    int64_t q = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    const StepCoef& sc = coef[blk_inst[q / 36]];
    Ktan[q] = (c1 * K[q] + sc.gbdt * C[q]) + sc.inv * M[q];
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) combine_tangent(int64_t nnz, const double* K, const double* C, const double* M,
                                                       double* Ktan, double c1, const StepCoef* __restrict__ coef,
                                                       const int32_t* __restrict__ blk_inst) { combine_tangent_b(nnz, K, C, M, Ktan, c1, coef, blk_inst, EB_VB); }
#endif

EBD void residual_corrector_b(int64_t ndofs, double alpha, const double* fext,
                                                          const double* Tc, const double* Tint, const double* fext_n,
                                                          const double* Tc_n, const double* Tint_n, const double* Tk,
                                                          double* res, const VB& vb) {
    int64_t d = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (d < ndofs)
        res[d] = ((1 + alpha) * ((fext[d] + Tc[d]) - Tint[d]) - alpha * ((fext_n[d] + Tc_n[d]) - Tint_n[d])) - Tk[d];
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) residual_corrector(int64_t ndofs, double alpha, const double* fext,
                                                          const double* Tc, const double* Tint, const double* fext_n,
                                                          const double* Tc_n, const double* Tint_n, const double* Tk,
                                                          double* res) { residual_corrector_b(ndofs, alpha, fext, Tc, Tint, fext_n, Tc_n, Tint_n, Tk, res, EB_VB); }
#endif

// predictor residual: r = (1+α)f - Tⁱⁿᵗⁿ - Tᵏⁿ - αfⁿ + Cⁿ·t₁ + (1/(βΔt²))·Mⁿ·t₂, the two products
// accumulated row-wise in increasing column order (what the CSC column sweep of mul! yields per row)
EBD void residual_predictor_b(int64_t ndofs, const int32_t* __restrict__ bptr,
                                                          const int32_t* __restrict__ bcol, double alpha,
                                                          const StepCoef* __restrict__ coef,
                                                          const int32_t* __restrict__ node_inst, const double* fext,
                                                          const double* Tint_n, const double* Tk_n,
                                                          const double* fext_n, const double* Dd, const double* Ddd,
                                                          const double* Cn, const double* Mn, double* res, const VB& vb) {
    int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    const int64_t n = i / 6;
    const int a = (int)(i % 6);
    const StepCoef& sc = coef[node_inst[n]];
    const double t1c = sc.gob, t2c = sc.t2c, dt = sc.dt, c3 = sc.inv;
    double r = (((1 + alpha) * fext[i] - Tint_n[i]) - Tk_n[i]) - alpha * fext_n[i];
    for (int b = bptr[n]; b < bptr[n + 1]; ++b) {
        const int64_t m = bcol[b];
        const double* row = Cn + (int64_t)36 * b + 6 * a;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            int64_t j = 6 * m + c;
            r += row[c] * (t1c * Dd[j] - t2c * Ddd[j]);
        }
    }
    for (int b = bptr[n]; b < bptr[n + 1]; ++b) {
        const int64_t m = bcol[b];
        const double* row = Mn + (int64_t)36 * b + 6 * a;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            int64_t j = 6 * m + c;
            r += row[c] * (c3 * (dt * Dd[j] + ((dt * dt) / 2) * Ddd[j]));
        }
    }
    res[i] = r;
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) residual_predictor(int64_t ndofs, const int32_t* __restrict__ bptr,
                                                          const int32_t* __restrict__ bcol, double alpha,
                                                          const StepCoef* __restrict__ coef,
                                                          const int32_t* __restrict__ node_inst, const double* fext,
                                                          const double* Tint_n, const double* Tk_n,
                                                          const double* fext_n, const double* Dd, const double* Ddd,
                                                          const double* Cn, const double* Mn, double* res) { residual_predictor_b(ndofs, bptr, bcol, alpha, coef, node_inst, fext, Tint_n, Tk_n, fext_n, Dd, Ddd, Cn, Mn, res, EB_VB); }
#endif

// apply_boundary_conditions! (utils_bcs_and_loads.jl:5-29): ΔD[d] = imposed - D[d] on fixed dofs
// and r -= ΔD[d]·Ktan[:, d], subtracted per row in increasing fixed-dof order
EBD void bc_lift_b(int64_t ndofs, const int32_t* __restrict__ bptr,
                                               const int32_t* __restrict__ bcol, const int32_t* __restrict__ fixidx,
                                               const double* imposed, const double* D, const double* Ktan, double* dD,
                                               double* res, Active act, const VB& vb) {
    int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    if (!act.flag[act.node_inst[i / 6]]) return;
    if (fixidx[i] >= 0) dD[i] = imposed[fixidx[i]] - D[i];
    const int64_t n = i / 6;
    const int a = (int)(i % 6);
    double r = res[i];
    bool any = false;
    for (int b = bptr[n]; b < bptr[n + 1]; ++b) {
        const int64_t m = bcol[b];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            int64_t j = 6 * m + c;
            int fj = fixidx[j];
            if (fj >= 0) {
                r -= (imposed[fj] - D[j]) * Ktan[(int64_t)36 * b + 6 * a + c];
                any = true;
            }
        }
    }
    if (any) res[i] = r;
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) bc_lift(int64_t ndofs, const int32_t* __restrict__ bptr,
                                               const int32_t* __restrict__ bcol, const int32_t* __restrict__ fixidx,
                                               const double* imposed, const double* D, const double* Ktan, double* dD,
                                               double* res, Active act) { bc_lift_b(ndofs, bptr, bcol, fixidx, imposed, D, Ktan, dD, res, act, EB_VB); }
#endif

// ---- masked block-row SpMV: y = A_ff x on free dofs, y = x on fixed dofs ---------------------
// (x is zero on fixed dofs throughout the Krylov iteration). Optional dot products with y.
// One workgroup per reduction chunk; chunks of instances whose solve is done are skipped.
struct KrylovCtl {
    double rho, rho_old, alpha, omega, bnorm2, rr2, ss2, tol2;
    int done, iters, breakdown;
    int pivot;     // 1: the instance's factorisation met a singular pivot block (its solve failed)
    int direct;    // 1: the chain preconditioner holds all the instance's couplings (direct first step)
};
// whether a preconditioner application of this solve works on the instance: its solve is open
// and, in the direct step (direct_only), its preconditioner is exact
EBD bool krylov_open(const KrylovCtl& k, int direct_only) { return !k.done && (!direct_only || k.direct); }

// row i of the solve's operator times x: the identity on fixed dofs, the tangent's block row
// (node i / 6, row i % 6) elsewhere, summed block by block in column order
EBD double masked_row(int64_t i, const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
                      const uint8_t* __restrict__ fixed, const double* __restrict__ A, const double* __restrict__ x) {
    if (fixed[i]) return x[i];
    const int64_t n = i / 6;
    const int a = (int)(i % 6);
    double acc = 0;
    for (int b = bptr[n]; b < bptr[n + 1]; ++b) {
        double rv[6], xv[6];
        ld6(A + (int64_t)36 * b + 6 * a, rv);
        ld6(x + 6 * (int64_t)bcol[b], xv);
#pragma unroll
        for (int c = 0; c < 6; ++c) acc += rv[c] * xv[c];
    }
    return acc;
}

template <int NQ>
EBD void spmv_masked_b(Segs sg, const int32_t* __restrict__ bptr,
                                                         const int32_t* __restrict__ bcol, const uint8_t* __restrict__ fixed,
                                                         const double* __restrict__ A, const double* __restrict__ x,
                                                         double* __restrict__ y, const double* __restrict__ dotw,
                                                         double* partials, const KrylovCtl* kc, const VB& vb) {
    if (kc[sg.inst[vb.x]].done) return;
    double v[NQ > 0 ? NQ : 1] = {0};
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK) {
        const double acc = masked_row(i, bptr, bcol, fixed, A, x);
        y[i] = acc;
        if (NQ >= 1) v[0] += dotw[i] * acc;
        if (NQ >= 2) v[NQ - 1] += acc * acc;
    }
    if (NQ > 0) seg_block_store<(NQ > 0 ? NQ : 1)>(v, partials, sg.nchunks, vb.x);
}
#ifndef EB_BODIES_ONLY
template <int NQ>
__global__ void __launch_bounds__(RED_BLOCK) spmv_masked(Segs sg, const int32_t* __restrict__ bptr,
                                                         const int32_t* __restrict__ bcol, const uint8_t* __restrict__ fixed,
                                                         const double* __restrict__ A, const double* __restrict__ x,
                                                         double* __restrict__ y, const double* __restrict__ dotw,
                                                         double* partials, const KrylovCtl* kc) { spmv_masked_b<NQ>(sg, bptr, bcol, fixed, A, x, y, dotw, partials, kc, EB_VB); }
#endif

// r = b − A x with the solve's operator (the identity on fixed dofs), on the chunks of unfinished
// solves: the right-hand side of the crossing chain cover's correction
EBD void residual_masked_b(Segs sg, const int32_t* __restrict__ bptr,
                                                             const int32_t* __restrict__ bcol,
                                                             const uint8_t* __restrict__ fixed,
                                                             const double* __restrict__ A, const double* __restrict__ x,
                                                             const double* __restrict__ b, double* __restrict__ r,
                                                             const KrylovCtl* kc, int direct_only, const VB& vb) {
    if (!krylov_open(kc[sg.inst[vb.x]], direct_only)) return;
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK) {
        const double acc = masked_row(i, bptr, bcol, fixed, A, x);
        r[i] = b[i] - acc;
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) residual_masked(Segs sg, const int32_t* __restrict__ bptr,
                                                             const int32_t* __restrict__ bcol,
                                                             const uint8_t* __restrict__ fixed,
                                                             const double* __restrict__ A, const double* __restrict__ x,
                                                             const double* __restrict__ b, double* __restrict__ r,
                                                             const KrylovCtl* kc, int direct_only) { residual_masked_b(sg, bptr, bcol, fixed, A, x, b, r, kc, direct_only, EB_VB); }
#endif

// ---- chain block-tridiagonal preconditioner ------------------------------------------------
// The node graph is covered by chains (paths); along a chain the masked tangent restricted to
// the chain is block-tridiagonal with 6x6 blocks (row i: L_i x_{i-1} + D_i x_i + U_i x_{i+1}) and
// is factored exactly by block cyclic reduction. Couplings between chains are left to the Krylov
// iteration.
//
// Cyclic reduction, one workgroup per chain. Level l (h = 2^l) eliminates the positions
// j = h, 3h, 5h, ... from the positions still active (the multiples of h):
//   phase A (eliminated j):  D_j <- D_j^-1,  GL_j = D_j^-1 L_j,  GU_j = D_j^-1 U_j
//   phase B (kept i = 0, 2h, 4h, ...; j1 = i-h, j2 = i+h):
//     D_i <- D_i - L_i GU_j1 - U_i GL_j2,   L_i <- -L_i GL_j1,   U_i <- -U_i GU_j2
// until position 0 is left alone and inverted. The L_i, U_i a kept position holds at each level
// are saved (hist) for the right-hand-side reduction of the solve. Depth: ceil(log2 n) levels
// instead of n sequential 6x6 steps, and every level is parallel over the chain.

constexpr int CHAIN_THREADS = 256;                       // factorisation workgroup
constexpr int CHAIN_GROUPS = (CHAIN_THREADS / 64) * 10;  // six-lane groups: ten per wavefront
constexpr int GJ_BUF = 24;                               // doubles of LDS exchange per group
// solve workgroup: 1024 threads while the chains are few (each chain's latency sets the time), 256
// once there are enough chains to fill the chip several times over (more chains per CU)
constexpr int SOLVE_THREADS_FEW = 1024, SOLVE_THREADS_MANY = 256;
constexpr int64_t SOLVE_MANY_CHAINS = 1024;

struct ChainDev {
    const int32_t* ptr;   // n_chains+1
    const int32_t* node;  // chain-ordered nodes
    const int32_t *diag_blk, *lo_blk, *up_blk;  // per chain position; -1 when absent
    const uint64_t *mD, *mL, *mU;  // bit k: entry k (row-major) touches a fixed dof row or column
    const int32_t* hbase;          // per chain position: first history slot
    double *Dm, *Lm, *Um;          // 36 per position: working blocks, then D^-1, GL, GU
    double* hist;                  // 72 per slot: L then U of a kept position at one level
    double *bw, *yw;               // 6 per position: solve work vectors (chains too long for LDS)
    const int32_t* inst;           // per chain: instance
    int64_t n_chains;
    const uint32_t* mrow;          // per position and row r (6 g + r): bits c, 6 + c, 12 + c: entry
                                   // (r, c) of D, L, U touches a fixed dof (mD, mL, mU restated)
};

EBD int chain_levels(int32_t n) {
    int nl = 0;
    while ((int64_t(1) << nl) < n) ++nl;
    return nl;
}

// A 6x6 block row-parallel: lane r of a six-lane group holds row r of [D | L | U] (NC = 18) or of
// D (NC = 6); the group shares GJ_BUF doubles of LDS for the pivot row. All six lanes of a group
// are in one wavefront, so the exchange needs no workgroup barrier, only the wavefront-scope
// ordering of the LDS accesses.
EBD void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// The same ordering of a wavefront's LDS accesses kept at compiler level only: the LDS executes a
// wavefront's instructions in issue order, so the lanes of a group see each other's writes. The
// fences of wave_sync also make the wavefront wait until its own global stores, and with them
// every load issued before those stores, have completed, which the block-LU kernels, whose loads
// run several steps ahead, must not do.
EBD void wave_lds_order() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}
template <bool F> EBD void wsync() {
    if (F) wave_sync();
    else wave_lds_order();
}

struct Lane6 {
    int group;  // six-lane group of the workgroup (lanes 60-63 of a wavefront take none)
    int r;      // row of the block this lane holds; 6 for the idle lanes
    double* buf;
};
EBD Lane6 lane6(double* bufs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Lane6 l;
    l.group = wave * 10 + (lane < 60 ? lane / 6 : 0);
    l.r = lane < 60 ? lane % 6 : 6;
    l.buf = bufs + GJ_BUF * l.group;
    return l;
}

// In-place Gauss-Jordan without row exchanges: on return the group holds row r of
// [D^-1 | D^-1 L | D^-1 U]. False (on the pivot row's lane) when its pivot has cancelled to below
// 1e-12 of the diagonal entry it started from, or is not finite; the caller then redoes the
// block with gj6_pivot. The test is invariant under row and column scaling, which matters here:
// translational and rotational entries differ by many decades.
template <int NC, bool F = true> EBD bool gj6_nopiv(double (&w)[NC], int r, double* buf) {
    double d0 = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) d0 = (r == k) ? fabs(w[k]) : d0;
    bool ok = true;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        if (r == p) {
            const double piv = w[p];
            ok = fabs(piv) > 1e-12 * d0 && fabs(piv) > 0.0 && fabs(piv) < INFINITY;
            const double ip = 1.0 / piv;
            w[p] = 1.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) w[c] *= ip;
#pragma unroll
            for (int c = 0; c < NC; c += 6) st6(buf + c, w + c);
        }
        wsync<F>();
        if (r != p) {
            double pr[NC];
#pragma unroll
            for (int c = 0; c < NC; c += 6) ld6(buf + c, pr + c);
            const double f = w[p];
            w[p] = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) w[c] = fma(-f, pr[c], w[c]);
        }
        wsync<F>();
    }
    return ok;
}

// Gauss-Jordan with partial pivoting on [D | I | L | U], the pivot of step p being the largest
// |entry| of column p among the rows not yet used (lowest row on ties). Rows are not moved: the
// lane whose row is the pivot of step p ends with row p of the result, returned in `orow`.
// False on that lane when its pivot is zero or not finite.
template <int NC, bool F = true> EBD bool gj6_pivot(double (&w)[NC], int r, double* buf, int& orow) {
    constexpr int NW = NC + 6;
    double x[NW];
#pragma unroll
    for (int c = 0; c < 6; ++c) { x[c] = w[c]; x[6 + c] = (r == c) ? 1.0 : 0.0; }
#pragma unroll
    for (int c = 6; c < NC; ++c) x[6 + c] = w[c];
    bool used = r >= 6, ok = true;
    orow = 0;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const double a = fabs(x[p]);
        if (r < 6) buf[r] = used ? -1.0 : (a >= 0.0 ? a : 0.0);
        wsync<F>();
        int q = 0;
        double best = -0.5;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double v = buf[k];
            if (v > best) { best = v; q = k; }
        }
        wsync<F>();
        if (r == q) {
            const double piv = x[p];
            ok = fabs(piv) > 0.0 && fabs(piv) < INFINITY;
            const double ip = 1.0 / piv;
#pragma unroll
            for (int c = 0; c < NW; ++c) x[c] *= ip;
            x[p] = 1.0;
#pragma unroll
            for (int c = 0; c < NW; c += 6) st6(buf + c, x + c);
            used = true;
            orow = p;
        }
        wsync<F>();
        if (r != q) {
            double pr[NW];
#pragma unroll
            for (int c = 0; c < NW; c += 6) ld6(buf + c, pr + c);
            const double f = x[p];
#pragma unroll
            for (int c = 0; c < NW; ++c) x[c] = fma(-f, pr[c], x[c]);
        }
        wsync<F>();
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) w[c] = x[6 + c];
    return ok;
}

// Inverse of the group's block (and D^-1 L, D^-1 U for NC = 18): unpivoted, redone with pivoting
// by the groups whose unpivoted elimination was rejected (load(w) fetches the rows again). `act`:
// the group holds a block this pass (uniform over the group). Every lane of the wavefront calls
// it. Returns the row of the result the lane holds.
// gj6_nopiv in one instruction stream for every lane of the group: the pivot row goes through
// LDS to all six lanes, and the pivot lane's scaling and the others' elimination become selects,
// so that no branch divides the wavefront and no register is copied where branches meet. The
// reciprocal of the pivot is v_rcp_f64 refined by two Newton steps (within a few ulp; the
// preconditioner does not need a correctly rounded division). Without fences (wave_lds_order).
template <int NC> EBD bool gj6_flat(double (&w)[NC], int r, double* buf) {
    double d0 = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) d0 = (r == k) ? fabs(w[k]) : d0;
    bool ok = true;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        if (r == p) {
#pragma unroll
            for (int c = 0; c < NC; c += 6) st6(buf + c, w + c);
        }
        wave_lds_order();
        double pr[NC];
#pragma unroll
        for (int c = 0; c < NC; c += 6) ld6(buf + c, pr + c);
        const double piv = pr[p];
        double ip = __builtin_amdgcn_rcp(piv);
        ip = fma(ip, fma(-piv, ip, 1.0), ip);
        ip = fma(ip, fma(-piv, ip, 1.0), ip);
        const bool pivot = r == p;
        ok = pivot ? (fabs(piv) > 1e-12 * d0 && fabs(piv) > 0.0 && fabs(piv) < INFINITY) : ok;
        const double f = w[p];
        pr[p] = 1.0;   // column p becomes the inverse's column
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double sc = pr[c] * ip;
            w[c] = pivot ? sc : fma(-f, sc, c == p ? 0.0 : w[c]);
        }
        wave_lds_order();
    }
    return ok;
}

template <int NC, typename Load, bool F = true>
EBD int gj6(double (&w)[NC], const Lane6& ln, bool act, bool& ok, const Load& load) {
    bool good;
    if constexpr (F) good = gj6_nopiv<NC, F>(w, ln.r, ln.buf);
    else good = gj6_flat<NC>(w, ln.r, ln.buf);
    const uint64_t bad = __ballot(act && ln.r < 6 && !good);
    const int lane = threadIdx.x & 63;
    const int first = lane < 60 ? lane - ln.r : 60;
    int orow = ln.r;
    if (act && ((bad >> first) & 0x3full)) {
        load(w);
        const bool nonsingular = gj6_pivot<NC, F>(w, ln.r, ln.buf, orow);
        if (ln.r < 6) ok = nonsingular && ok;
    }
    return orow;
}

// Masked entry k of position g's block (kind 0: D, 1: L, 2: U) read from the tangent: an entry in
// a fixed row or column becomes the identity's; an absent block is zero.
EBD double masked_entry(const ChainDev& cd, const double* __restrict__ A, int64_t g, int kind, int k) {
    const uint64_t m = kind == 0 ? cd.mD[g] : (kind == 1 ? cd.mL[g] : cd.mU[g]);
    const int32_t blk = kind == 0 ? cd.diag_blk[g] : (kind == 1 ? cd.lo_blk[g] : cd.up_blk[g]);
    if ((m >> k) & 1) return (kind == 0 && k / 6 == k % 6) ? 1.0 : 0.0;
    return blk >= 0 ? A[36 * (int64_t)blk + k] : 0.0;
}

// Where a level reads its blocks and writes the kept positions' new blocks: the masked tangent
// (first level only), the global working arrays, or the workgroup's LDS. Once positions are
// 2^L0 apart (levels l ≥ L0) the chain's active blocks live in LDS, at slot p >> L0; the factors
// the solve needs (D⁻¹, GL, GU at elimination, history) always go to the global arrays.
enum { CH_A = 0, CH_G = 1, CH_S = 2 };

struct ChainLDS {
    double *D, *L, *U;  // [slots][36]
    int L0;
};

template <int SRC>
EBD const double* chain_src(const ChainDev& cd, const ChainLDS& sm, int64_t g, int32_t p, int kind) {
    if (SRC == CH_S) return (kind == 0 ? sm.D : (kind == 1 ? sm.L : sm.U)) + 36 * (int64_t)(p >> sm.L0);
    return (kind == 0 ? cd.Dm : (kind == 1 ? cd.Lm : cd.Um)) + 36 * g;
}

template <int DST> EBD double* chain_dst(const ChainDev& cd, const ChainLDS& sm, int64_t g, int32_t p, int kind) {
    if (DST == CH_S) return (kind == 0 ? sm.D : (kind == 1 ? sm.L : sm.U)) + 36 * (int64_t)(p >> sm.L0);
    return (kind == 0 ? cd.Dm : (kind == 1 ? cd.Lm : cd.Um)) + 36 * g;
}

// row r of block `kind` of position p (g = s0 + p)
template <int SRC>
EBD void chain_row(const ChainDev& cd, const ChainLDS& sm, const double* __restrict__ A, int64_t g, int32_t p, int kind,
                   int r, double* out) {
    if (SRC == CH_A) {
        const uint64_t m = kind == 0 ? cd.mD[g] : (kind == 1 ? cd.mL[g] : cd.mU[g]);
        const int32_t blk = kind == 0 ? cd.diag_blk[g] : (kind == 1 ? cd.lo_blk[g] : cd.up_blk[g]);
        double v[6];
        ld6(A + 36 * (int64_t)(blk >= 0 ? blk : 0) + 6 * r, v);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            out[c] = ((m >> (6 * r + c)) & 1) ? ((kind == 0 && r == c) ? 1.0 : 0.0) : (blk >= 0 ? v[c] : 0.0);
    } else {
        ld6(chain_src<SRC>(cd, sm, g, p, kind) + 6 * r, out);
    }
}

// acc += row · G (6x6 row-major), the block's loads fenced off from the neighbouring code
EBD void row_times_block(const double (&row)[6], const double* __restrict__ G, double (&acc)[6]) {
    double g[36];
#pragma unroll
    for (int k = 0; k < 36; k += 6) ld6(G + k, g + k);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[c] = fma(row[k], g[6 * k + c], acc[c]);
    }
    __builtin_amdgcn_sched_barrier(0);
}

// one reduction level: SRC where its blocks are, DST where the kept positions' new blocks go
template <int SRC, int DST>
EBD void chain_level(const ChainDev& cd, const ChainLDS& sm, const double* __restrict__ A, int64_t s0, int32_t n, int l,
                     bool& ok, const Lane6& ln) {
    const int32_t h = 1 << l;
    // phase A: a six-lane group per eliminated position j = h, 3h, 5h, ...:
    //   D_j <- D_j^-1, GL_j = D_j^-1 L_j, GU_j = D_j^-1 U_j  (one Gauss-Jordan on [D | L | U])
    const int32_t ne = (n - h + 2 * h - 1) / (2 * h);
    for (int32_t base = 0; base < ne; base += CHAIN_GROUPS) {
        const int32_t m = base + ln.group;
        const bool act = m < ne;
        const int32_t p = act ? h + 2 * h * m : 0;
        const int64_t g = s0 + p;
        const int rr = ln.r < 6 ? ln.r : 0;
        auto load = [&](double (&w)[18]) {
            if (act) {
                chain_row<SRC>(cd, sm, A, g, p, 0, rr, w);
                chain_row<SRC>(cd, sm, A, g, p, 1, rr, w + 6);
                chain_row<SRC>(cd, sm, A, g, p, 2, rr, w + 12);
            } else {
#pragma unroll
                for (int c = 0; c < 18; ++c) w[c] = 0.0;
            }
        };
        double w[18];
        load(w);
        const int orow = gj6<18>(w, ln, act, ok, load);
        if (act && ln.r < 6) {
            const int64_t o = 36 * g + 6 * orow;
            st6(cd.Dm + o, w); st6(cd.Lm + o, w + 6); st6(cd.Um + o, w + 12);
            if (SRC == CH_S) {
                const int64_t so = 36 * (int64_t)(p >> sm.L0) + 6 * orow;
                st6(sm.D + so, w); st6(sm.L + so, w + 6); st6(sm.U + so, w + 12);
            }
        }
    }
    __syncthreads();
    // phase B: a six-lane group per kept position i = 0, 2h, 4h, ... (j1 = i-h, j2 = i+h), lane r
    // forming row r of
    //   D_i <- D_i - L_i GU_j1 - U_i GL_j2,   L_i <- -L_i GL_j1,   U_i <- -U_i GU_j2
    // The L_i, U_i it started from are saved (history) for the solve's right-hand-side reduction.
    // The eliminated neighbours' blocks come from where phase A put them.
    constexpr int NSRC = SRC == CH_S ? CH_S : CH_G;
    const int32_t nk = (n + 2 * h - 1) / (2 * h);
    for (int32_t base = 0; base < nk; base += CHAIN_GROUPS) {
        const int32_t m = base + ln.group;
        if (m >= nk || ln.r >= 6) continue;
        const int r = ln.r;
        const int32_t i = 2 * h * m;
        const int64_t gi = s0 + i;
        const bool left = i > 0, right = i + h < n;
        double Lr[6], Ur[6], Dr[6];
        chain_row<SRC>(cd, sm, A, gi, i, 0, r, Dr);
        chain_row<SRC>(cd, sm, A, gi, i, 1, r, Lr);
        chain_row<SRC>(cd, sm, A, gi, i, 2, r, Ur);
        // one neighbour block at a time (its 36 entries in flight together), so that the lane's
        // registers hold one block rather than four
        double a1[6] = {}, a2[6] = {}, nl[6] = {}, nu[6] = {};
        if (left) {
            row_times_block(Lr, chain_src<NSRC>(cd, sm, gi - h, i - h, 2), a1);
            row_times_block(Lr, chain_src<NSRC>(cd, sm, gi - h, i - h, 1), nl);
        }
        if (right) {
            row_times_block(Ur, chain_src<NSRC>(cd, sm, gi + h, i + h, 1), a2);
            row_times_block(Ur, chain_src<NSRC>(cd, sm, gi + h, i + h, 2), nu);
        }
        double* hs = cd.hist + 72 * ((int64_t)cd.hbase[gi] + l) + 6 * r;
        if (left) st6(hs, Lr);
        if (right) st6(hs + 36, Ur);
        // a missing neighbour leaves a zero coupling block (written where the blocks now live)
        double t[6];
        if (left || SRC == CH_A || DST != SRC) {
#pragma unroll
            for (int c = 0; c < 6; ++c) t[c] = left ? -nl[c] : 0.0;
            st6(chain_dst<DST>(cd, sm, gi, i, 1) + 6 * r, t);
        }
        if (right || SRC == CH_A || DST != SRC) {
#pragma unroll
            for (int c = 0; c < 6; ++c) t[c] = right ? -nu[c] : 0.0;
            st6(chain_dst<DST>(cd, sm, gi, i, 2) + 6 * r, t);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) t[c] = (Dr[c] - a1[c]) - a2[c];
        st6(chain_dst<DST>(cd, sm, gi, i, 0) + 6 * r, t);

    }
    __syncthreads();
}

// Cyclic reduction of each chain's block-tridiagonal system, one workgroup per chain.
// Dynamic LDS: 3 × 36 doubles per slot, ceil(n / 2^L0) slots (chain_lds_bytes).
// status[instance] is set to 1 when a pivot block of one of its chains is singular (solve_begin
// clears the flags before each solve).
// Two wavefronts per SIMD (at most 256 registers), so that two chains share a CU: a launch has
// about 1.5 chains per CU.
EBD void chain_factor_b(ChainDev cd, const double* __restrict__ A, int* status, int L0, const uint8_t* __restrict__ active, const VB& vb) {
    extern __shared__ __align__(16) double chain_smem[];
    __shared__ __align__(16) double gj_buf[CHAIN_GROUPS * GJ_BUF];
    if (!active[cd.inst[vb.x]]) return;   // instance frozen in the Newton iteration
    const int64_t s0 = cd.ptr[vb.x];
    const int32_t n = cd.ptr[vb.x + 1] - (int32_t)s0;
    const int32_t slots = (n + (1 << L0) - 1) >> L0;
    const ChainLDS sm{chain_smem, chain_smem + 36 * (int64_t)slots, chain_smem + 72 * (int64_t)slots, L0};
    const Lane6 ln = lane6(gj_buf);
    bool ok = true;
    const int nl = chain_levels(n);
    for (int l = 0; l < nl; ++l) {
        const bool to_lds = l + 1 >= L0;
        if (l == 0) {
            if (to_lds) chain_level<CH_A, CH_S>(cd, sm, A, s0, n, l, ok, ln);
            else chain_level<CH_A, CH_G>(cd, sm, A, s0, n, l, ok, ln);
        } else if (l < L0) {
            if (to_lds) chain_level<CH_G, CH_S>(cd, sm, A, s0, n, l, ok, ln);
            else chain_level<CH_G, CH_G>(cd, sm, A, s0, n, l, ok, ln);
        } else {
            chain_level<CH_S, CH_S>(cd, sm, A, s0, n, l, ok, ln);
        }
    }
    // position 0, left alone: its D inverted by the first group
    if (threadIdx.x < 64) {
        const bool act = ln.group == 0;
        const int rr = ln.r < 6 ? ln.r : 0;
        auto load = [&](double (&w)[6]) {
            if (nl == 0) chain_row<CH_A>(cd, sm, A, s0, 0, 0, rr, w);
            else if (nl >= L0) chain_row<CH_S>(cd, sm, A, s0, 0, 0, rr, w);
            else chain_row<CH_G>(cd, sm, A, s0, 0, 0, rr, w);
        };
        double w[6];
        load(w);
        const int orow = gj6<6>(w, ln, act, ok, load);
        if (act && ln.r < 6) st6(cd.Dm + 36 * s0 + 6 * orow, w);
    }
    if (!ok) status[cd.inst[vb.x]] = 1;
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(CHAIN_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
chain_factor(ChainDev cd, const double* __restrict__ A, int* status, int L0, const uint8_t* __restrict__ active) { chain_factor_b(cd, A, status, L0, active, EB_VB); }
#endif

// smallest L0 whose LDS share for the longest chain stays within `budget` bytes
inline int chain_lds_level(int64_t max_len, int64_t budget) {
    int L0 = 0;
    while (((max_len + (int64_t(1) << L0) - 1) >> L0) * 3 * 36 * (int64_t)sizeof(double) > budget) ++L0;
    return L0;
}
inline size_t chain_lds_bytes(int64_t max_len, int L0) {
    return (size_t)(((max_len + (int64_t(1) << L0) - 1) >> L0) * 3 * 36 * (int64_t)sizeof(double));
}
// the solve keeps its two work vectors (12 doubles per position) in LDS while they fit in `cap`
// bytes (what the device grants a workgroup beside the kernel's static LDS); 0: global memory
inline size_t chain_solve_lds_bytes(int64_t max_len, int64_t cap) {
    const int64_t b = 12 * max_len * (int64_t)sizeof(double);
    return b <= cap ? (size_t)b : 0;
}

// x = M^-1 b along each chain, with the factors of chain_factor: the right-hand side is reduced
// level by level (y_j = D_j^-1 b_j for eliminated j; b_i -= L_i y_j1 + U_i y_j2 for kept i with the
// saved blocks), then solved back from position 0 (x_j = y_j - GL_j x_{j-h} - GU_j x_{j+h}).
// Six lanes per position, one per row; the work vectors in LDS (LDSV) or in the global arrays.
// The solve of one chain (s0, n) on its right-hand side in bw (6 per position, overwritten): the
// reduction level by level, the root, the back-substitution; the solution ends in yw. Every
// thread of the workgroup calls it (it holds workgroup barriers).
template <int NT>
EBD void chain_solve_body(const ChainDev& cd, int64_t s0, int32_t n, double* bw, double* yw) {
    const int tid = threadIdx.x;
    constexpr int SOLVE_THREADS = NT;
    const int nl = chain_levels(n);
    for (int l = 0; l < nl; ++l) {
        const int32_t h = 1 << l;
        const int32_t ne = (n - h + 2 * h - 1) / (2 * h);
        for (int32_t t = tid; t < 6 * ne; t += SOLVE_THREADS) {
            const int32_t p = h + 2 * h * (t / 6);
            const int r = t % 6;
            double Dr[6], bv[6];
            ld6(cd.Dm + 36 * (s0 + p) + 6 * r, Dr);
            ld6(bw + 6 * p, bv);
            double acc = 0;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc = fma(Dr[c], bv[c], acc);
            yw[6 * p + r] = acc;
        }
        __syncthreads();
        const int32_t nk = (n + 2 * h - 1) / (2 * h);
        for (int32_t t = tid; t < 6 * nk; t += SOLVE_THREADS) {
            const int32_t i = 2 * h * (t / 6);
            const int r = t % 6;
            const double* hs = cd.hist + 72 * ((int64_t)cd.hbase[s0 + i] + l) + 6 * r;
            double v = bw[6 * i + r];
            if (i > 0) {
                double hr[6], yv[6];
                ld6(hs, hr);
                ld6(yw + 6 * (i - h), yv);
#pragma unroll
                for (int c = 0; c < 6; ++c) v = fma(-hr[c], yv[c], v);
            }
            if (i + h < n) {
                double hr[6], yv[6];
                ld6(hs + 36, hr);
                ld6(yw + 6 * (i + h), yv);
#pragma unroll
                for (int c = 0; c < 6; ++c) v = fma(-hr[c], yv[c], v);
            }
            bw[6 * i + r] = v;
        }
        __syncthreads();
    }
    if (tid < 6) {
        double Dr[6], bv[6];
        ld6(cd.Dm + 36 * s0 + 6 * tid, Dr);
        ld6(bw, bv);
        double acc = 0;
#pragma unroll
        for (int c = 0; c < 6; ++c) acc = fma(Dr[c], bv[c], acc);
        yw[tid] = acc;
    }
    __syncthreads();
    for (int l = nl - 1; l >= 0; --l) {
        const int32_t h = 1 << l;
        const int32_t ne = (n - h + 2 * h - 1) / (2 * h);
        for (int32_t t = tid; t < 6 * ne; t += SOLVE_THREADS) {
            const int32_t j = h + 2 * h * (t / 6);
            const int r = t % 6;
            const int64_t g = s0 + j;
            double v = yw[6 * j + r];
            {
                double gr[6], yv[6];
                ld6(cd.Lm + 36 * g + 6 * r, gr);
                ld6(yw + 6 * (j - h), yv);
#pragma unroll
                for (int c = 0; c < 6; ++c) v = fma(-gr[c], yv[c], v);
            }
            if (j + h < n) {
                double gr[6], yv[6];
                ld6(cd.Um + 36 * g + 6 * r, gr);
                ld6(yw + 6 * (j + h), yv);
#pragma unroll
                for (int c = 0; c < 6; ++c) v = fma(-gr[c], yv[c], v);
            }
            yw[6 * j + r] = v;
        }
        __syncthreads();
    }
}

// ACCUM: x += M⁻¹ b (the crossing cover's correction) instead of x = M⁻¹ b
template <bool LDSV, int SOLVE_THREADS, bool ACCUM = false>
EBD void chain_solve_b(ChainDev cd, const double* __restrict__ b,
                                                             double* __restrict__ x, const KrylovCtl* kc,
                                                             int direct_only, const VB& vb) {
    extern __shared__ __align__(16) double solve_smem[];
    if (!krylov_open(kc[cd.inst[vb.x]], direct_only)) return;
    const int64_t s0 = cd.ptr[vb.x];
    const int32_t n = cd.ptr[vb.x + 1] - (int32_t)s0;
    const int tid = threadIdx.x;
    double* bw = LDSV ? solve_smem : cd.bw + 6 * s0;
    double* yw = LDSV ? solve_smem + 6 * (int64_t)n : cd.yw + 6 * s0;
    for (int32_t t = tid; t < 6 * n; t += SOLVE_THREADS) bw[t] = b[6 * (int64_t)cd.node[s0 + t / 6] + t % 6];
    __syncthreads();
    chain_solve_body<SOLVE_THREADS>(cd, s0, n, bw, yw);
    for (int32_t t = tid; t < 6 * n; t += SOLVE_THREADS) {
        double& xi = x[6 * (int64_t)cd.node[s0 + t / 6] + t % 6];
        xi = ACCUM ? xi + yw[t] : yw[t];
    }
}
#ifndef EB_BODIES_ONLY
template <bool LDSV, int SOLVE_THREADS, bool ACCUM = false>
__global__ void __launch_bounds__(SOLVE_THREADS) chain_solve(ChainDev cd, const double* __restrict__ b,
                                                             double* __restrict__ x, const KrylovCtl* kc,
                                                             int direct_only) { chain_solve_b<LDSV, SOLVE_THREADS, ACCUM>(cd, b, x, kc, direct_only, EB_VB); }
#endif

// ---- block LU along each chain, from both ends ----------------------------------------------
// The same block-tridiagonal system factored position by position (block Thomas), from both ends
// of the chain at once and meeting in the middle, m = (n − 1) / 2:
//   top    (p = 0 .. m−1):   D'_p = D_p − L_p G_{p−1},   G_p = D'_p⁻¹ U_p
//   bottom (p = n−1 .. m+1): D'_p = D_p − U_p H_{p+1},   H_p = D'_p⁻¹ L_p
//   middle:                  D'_m = D_m − L_m G_{m−1} − U_m H_{m+1}
// kept as Dm = D'_p⁻¹ and, above the middle, Um = G_p, Lm = L_p, below it Lm = H_p, Um = U_p
// (at m: Lm = L_m, Um = U_m), and solved by
//   y_p = D'_p⁻¹ (b_p − L_p y_{p−1}) (top),   y_p = D'_p⁻¹ (b_p − U_p y_{p+1}) (bottom),
//   x_m = D'_m⁻¹ (b_m − L_m y_{m−1} − U_m y_{m+1}),
//   x_p = y_p − G_p x_{p+1} (top),   x_p = y_p − H_p x_{p−1} (bottom).
// About n/2 dependent steps per chain instead of cyclic reduction's ceil(log2 n) levels, but half
// its flops, no workgroup barrier and no level history: the form for batches of many chains,
// where the chains, not the positions of one chain, fill the chip. One six-lane group per end
// (lane r holds row r), the two groups of a chain in one wavefront, five chains per 64-lane
// workgroup.
//
// Loads run in two stages ahead of the step that uses them: the indices of the step LU_AHEAD
// (the factorisation: LU_FACTOR_AHEAD) further on (block numbers, fixed-dof masks, node numbers)
// two stages ahead, the data at those indices one stage ahead. No load waits on a load issued in
// the same step, so a step waits only for data requested that many steps earlier. Steps past the
// end of a half read position 0 (any valid position): they are inactive and store nothing. The
// wavefront runs its longest half's steps, rounded up to a multiple of that distance; the step body
// has no branch but masked stores.
constexpr int LU_GROUPS = 10;
constexpr int LU_CHAINS = LU_GROUPS / 2;
constexpr int LU_AHEAD = 4;
// the factorisation's distance: with two steps of blocks in flight it needs at most 256 VGPRs per
// lane, so a SIMD holds two of its wavefronts (with four steps: 358 VGPRs, one wavefront), which
// on a batch of many chains outweighs the longer waits
constexpr int LU_FACTOR_AHEAD = 2;

EBD int32_t wave_max_i32(int32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// the lane's half of its chain: group 2c + h of the workgroup holds chain c's half h (0: top)
struct LuHalf {
    int64_t s0;      // the chain's first position
    int32_t n, m;    // positions, middle
    int32_t steps;   // steps of this half (0 when the chain takes no part)
    bool top;
    EBD int32_t at(int32_t q) const {   // position of step q, any valid position past the end
        return q < steps ? (top ? q : n - 1 - q) : 0;
    }
};
EBD LuHalf lu_half(const ChainDev& cd, int64_t ch, bool take, int group) {
    LuHalf h{0, 0, 0, 0, (group & 1) == 0};
    if (take) {
        h.s0 = cd.ptr[ch];
        h.n = cd.ptr[ch + 1] - (int32_t)h.s0;
        h.m = (h.n - 1) / 2;
        h.steps = h.top ? h.m : h.n - 1 - h.m;
    }
    return h;
}

// block numbers of D_p, L_p, U_p and the fixed-dof bits of their row r
struct LuIx {
    int32_t b[3];
    uint32_t m;
};
EBD void lu_ix(const ChainDev& cd, int64_t g, int r, LuIx& ix) {
    ix.b[0] = cd.diag_blk[g];
    ix.b[1] = cd.lo_blk[g];
    ix.b[2] = cd.up_blk[g];
    ix.m = cd.mrow[6 * g + r];
}
// row r of the three blocks as the tangent holds them (an absent block reads block 0: masked off)
EBD void lu_raw(const double* __restrict__ A, const LuIx& ix, int r, double (&v)[18]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) ld6(A + 36 * (int64_t)(ix.b[k] >= 0 ? ix.b[k] : 0) + 6 * r, v + 6 * k);
}
// the masked rows: an entry in a fixed row or column becomes the identity's, an absent block zero
EBD void lu_mask(const LuIx& ix, int r, double (&v)[18]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
            v[6 * k + c] = ((ix.m >> (6 * k + c)) & 1) ? ((k == 0 && r == c) ? 1.0 : 0.0)
                                                       : (ix.b[k] >= 0 ? v[6 * k + c] : 0.0);
    }
}
// w −= row · G (6x6 row-major in LDS)
EBD void lu_sub_row_block(double* w, const double (&row)[6], const double* G) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double gj[6];
        ld6(G + 6 * j, gj);
#pragma unroll
        for (int c = 0; c < 6; ++c) w[c] = fma(-row[j], gj[c], w[c]);
    }
}

// status[instance] is set to 1 when a pivot block of one of its chains is singular
EBD void chain_lu_factor_b(ChainDev cd, const double* __restrict__ A, int* status,
                                                      const uint8_t* __restrict__ active, const VB& vb) {
    __shared__ __align__(16) double gj_buf[LU_GROUPS * GJ_BUF];
    __shared__ __align__(16) double g_prev[LU_GROUPS * 36];
    const Lane6 ln = lane6(gj_buf);
    const int64_t ch = (int64_t)vb.x * LU_CHAINS + (ln.group >> 1);
    const bool mine = ln.r < 6 && ch < cd.n_chains;
    const bool on = mine && active[cd.inst[ch]];
    const LuHalf h = lu_half(cd, ch, on, ln.group);
    const int32_t smax = wave_max_i32(h.steps);
    const int r = ln.r < 6 ? ln.r : 0;
    double* gp = g_prev + 36 * ln.group;
    // G_{−1} = H_n = 0, so that the first step of each half is like the others
    if (ln.r < 6) {
        const double z[6] = {0, 0, 0, 0, 0, 0};
        st6(gp + 6 * ln.r, z);
    }
    wave_lds_order();
    bool ok = true;
    LuIx far[LU_FACTOR_AHEAD], near[LU_FACTOR_AHEAD];
    double raw[LU_FACTOR_AHEAD][18];
    // the first steps' loads in the order of the loop's own (data of step k, then the indices of
    // step k + LU_FACTOR_AHEAD), so that the loop starts with the waits of its steady state
#pragma unroll
    for (int k = 0; k < LU_FACTOR_AHEAD; ++k) lu_ix(cd, h.s0 + h.at(k), r, near[k]);
#pragma unroll
    for (int k = 0; k < LU_FACTOR_AHEAD; ++k) {
        lu_raw(A, near[k], r, raw[k]);
        lu_ix(cd, h.s0 + h.at(LU_FACTOR_AHEAD + k), r, far[k]);
    }
    for (int32_t q0 = 0; q0 < smax; q0 += LU_FACTOR_AHEAD) {
#pragma unroll
        for (int k = 0; k < LU_FACTOR_AHEAD; ++k) {
            const int32_t q = q0 + k;
            double v[18];
#pragma unroll
            for (int c = 0; c < 18; ++c) v[c] = raw[k][c];
            const LuIx ix = near[k];
            near[k] = far[k];
            lu_raw(A, near[k], r, raw[k]);                        // step q + LU_FACTOR_AHEAD
            lu_ix(cd, h.s0 + h.at(q + 2 * LU_FACTOR_AHEAD), r, far[k]);     // step q + 2 LU_FACTOR_AHEAD
            lu_mask(ix, r, v);
            const bool act = q < h.steps;
            // the block towards the finished side (lo) and towards the middle (up)
            double lo[6], up[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                lo[c] = h.top ? v[6 + c] : v[12 + c];
                up[c] = h.top ? v[12 + c] : v[6 + c];
            }
            // row r of [D'_p | up], D'_p = D_p − lo · (the previous step's G or H, in LDS)
            double w[12];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                w[c] = v[c];
                w[6 + c] = up[c];
            }
            lu_sub_row_block(w, lo, gp);
            double w0[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) w0[c] = w[c];
            auto reload = [&](double (&x)[12]) {
#pragma unroll
                for (int c = 0; c < 12; ++c) x[c] = w0[c];
            };
            const int orow = gj6<12, decltype(reload), false>(w, ln, act, ok, reload);
            if (act) {
                const int64_t g = h.s0 + h.at(q);
                st6(cd.Dm + 36 * g + 6 * orow, w);
                st6((h.top ? cd.Um : cd.Lm) + 36 * g + 6 * orow, w + 6);
                st6((h.top ? cd.Lm : cd.Um) + 36 * g + 6 * r, lo);
                st6(gp + 6 * orow, w + 6);
            }
            wave_lds_order();
        }
    }
    // the middle, by the top group: D'_m = D_m − L_m G_{m−1} − U_m H_{m+1}
    {
        const bool act = on && h.top;
        const int64_t g = h.s0 + (act ? h.m : 0);
        LuIx ix;
        lu_ix(cd, g, r, ix);
        double v[18];
        lu_raw(A, ix, r, v);
        lu_mask(ix, r, v);
        double w[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) w[c] = v[c];
        double Lr[6], Ur[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            Lr[c] = v[6 + c];
            Ur[c] = v[12 + c];
        }
        lu_sub_row_block(w, Lr, gp);        // G_{m−1} (zero when the top half is empty)
        lu_sub_row_block(w, Ur, gp + 36);   // H_{m+1}, the bottom group's (zero when it is empty)
        double w0[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) w0[c] = w[c];
        auto reload = [&](double (&x)[6]) {
#pragma unroll
            for (int c = 0; c < 6; ++c) x[c] = w0[c];
        };
        const int orow = gj6<6, decltype(reload), false>(w, ln, act, ok, reload);
        if (act) {
            st6(cd.Dm + 36 * g + 6 * orow, w);
            st6(cd.Lm + 36 * g + 6 * r, Lr);
            st6(cd.Um + 36 * g + 6 * r, Ur);
        }
    }
    if (mine && !ok) status[cd.inst[ch]] = 1;
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(64) chain_lu_factor(ChainDev cd, const double* __restrict__ A, int* status,
                                                      const uint8_t* __restrict__ active) { chain_lu_factor_b(cd, A, status, active, EB_VB); }
#endif

// The solve of the lane's half (see LuHalf; smax: the wavefront's longest half) with the factors
// of chain_lu_factor. io.node(p) is the index io.b and io.put address position p by (loaded two
// stages ahead), io.b(node, p) row r of b_p, io.old(node, p) what io.put(node, p, old, x, w) adds
// row r of x_p to (loaded one stage ahead); w points at position p's row r in wk, which holds 6
// doubles per position (y). v6: the group's 6-double LDS exchange, t6 its scratch; the two groups
// of a chain hold theirs side by side (the top group's first). Every lane of the wavefront calls it;
// `live`: the lane holds a row of a participating chain.
template <class IO>
EBD void chain_lu_solve_group(const ChainDev& cd, const LuHalf& h, int32_t smax, int r, bool live, double* wk,
                              double* v6, double* t6, const IO& io) {
    const double* lo_m = h.top ? cd.Lm : cd.Um;   // the block towards the finished side
    const double* up_m = h.top ? cd.Um : cd.Lm;   // G or H
    if (live) v6[r] = 0.0;                        // y_{−1} = y_n = 0
    wave_lds_order();
    // forward: y_p = D'_p⁻¹ (b_p − lo_p y_prev)
    {
        int64_t far[LU_AHEAD];
        double L[LU_AHEAD][6], D[LU_AHEAD][6], bb[LU_AHEAD];
        auto near_load = [&](int k, int32_t q, int64_t node) {
            const int32_t p = h.at(q);
            ld6(lo_m + 36 * (h.s0 + p) + 6 * r, L[k]);
            ld6(cd.Dm + 36 * (h.s0 + p) + 6 * r, D[k]);
            bb[k] = io.b(node, p);
        };
        // the first steps' loads in the loop's own order (see chain_lu_factor)
        int64_t first[LU_AHEAD];
#pragma unroll
        for (int k = 0; k < LU_AHEAD; ++k) first[k] = io.node(h.at(k));
#pragma unroll
        for (int k = 0; k < LU_AHEAD; ++k) {
            near_load(k, k, first[k]);
            far[k] = io.node(h.at(LU_AHEAD + k));
        }
        for (int32_t q0 = 0; q0 < smax; q0 += LU_AHEAD) {
#pragma unroll
            for (int k = 0; k < LU_AHEAD; ++k) {
                const int32_t q = q0 + k;
                double cL[6], cD[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    cL[c] = L[k][c];
                    cD[c] = D[k][c];
                }
                double t = bb[k];
                {
                    double yv[6];
                    ld6(v6, yv);
#pragma unroll
                    for (int c = 0; c < 6; ++c) t = fma(-cL[c], yv[c], t);
                }
                if (live) t6[r] = t;
                wave_lds_order();
                double tv[6];
                ld6(t6, tv);
                double y = 0.0;
#pragma unroll
                for (int c = 0; c < 6; ++c) y = fma(cD[c], tv[c], y);
                wave_lds_order();
                if (live && q < h.steps) {   // past its half's end a group keeps its last y
                    wk[6 * h.at(q) + r] = y;
                    v6[r] = y;
                }
                wave_lds_order();
                // the slot's refill once its values are used: the loads go to the same registers
                near_load(k, q + LU_AHEAD, far[k]);
                far[k] = io.node(h.at(q + 2 * LU_AHEAD));
            }
        }
    }
    // the middle, by the top group: x_m = D'_m⁻¹ (b_m − L_m y_{m−1} − U_m y_{m+1}); then both
    // groups start their back-substitution from x_m
    {
        const bool act = live && h.top;
        const int32_t m = act ? h.m : 0;
        const int64_t node = io.node(m);
        double Lr[6], Ur[6], Dr[6];
        ld6(cd.Lm + 36 * (h.s0 + m) + 6 * r, Lr);
        ld6(cd.Um + 36 * (h.s0 + m) + 6 * r, Ur);
        ld6(cd.Dm + 36 * (h.s0 + m) + 6 * r, Dr);
        double t = io.b(node, m);
        const double xo = io.old(node, m);
        double ya[6], yb[6];
        ld6(v6, ya);        // y_{m−1} (top group)
        ld6(v6 + 12, yb);   // y_{m+1} (its bottom group, 12 doubles on)
#pragma unroll
        for (int c = 0; c < 6; ++c) t = fma(-Lr[c], ya[c], fma(-Ur[c], yb[c], t));
        if (act) t6[r] = t;
        wave_lds_order();
        double tv[6];
        ld6(t6, tv);
        double x = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) x = fma(Dr[c], tv[c], x);
        wave_lds_order();
        if (act) {
            v6[r] = x;
            v6[12 + r] = x;
            io.put(node, m, xo, x, wk + 6 * m + r);
        }
        wave_lds_order();
    }
    // backward: x_p = y_p − up_p x_next, from the middle outwards (step q of the half: q counted
    // from the middle, the reverse of the forward order)
    {
        auto bat = [&](int32_t q) { return h.at(h.steps - 1 - q); };   // valid for q < steps
        auto bpos = [&](int32_t q) { return q < h.steps ? bat(q) : 0; };
        int64_t far[LU_AHEAD], nd[LU_AHEAD];
        double G[LU_AHEAD][6], yy[LU_AHEAD], old[LU_AHEAD];
        auto near_load = [&](int k, int32_t q, int64_t node) {
            const int32_t p = bpos(q);
            ld6(up_m + 36 * (h.s0 + p) + 6 * r, G[k]);
            yy[k] = wk[6 * p + r];
            old[k] = io.old(node, p);
            nd[k] = node;
        };
        int64_t first[LU_AHEAD];
#pragma unroll
        for (int k = 0; k < LU_AHEAD; ++k) first[k] = io.node(bpos(k));
#pragma unroll
        for (int k = 0; k < LU_AHEAD; ++k) {
            near_load(k, k, first[k]);
            far[k] = io.node(bpos(LU_AHEAD + k));
        }
        for (int32_t q0 = 0; q0 < smax; q0 += LU_AHEAD) {
#pragma unroll
            for (int k = 0; k < LU_AHEAD; ++k) {
                const int32_t q = q0 + k;
                double cG[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) cG[c] = G[k][c];
                double x = yy[k];
                const double xo = old[k];
                const int64_t node = nd[k];
                {
                    double xv[6];
                    ld6(v6, xv);
#pragma unroll
                    for (int c = 0; c < 6; ++c) x = fma(-cG[c], xv[c], x);
                }
                wave_lds_order();
                if (live && q < h.steps) {
                    const int32_t p = bat(q);
                    v6[r] = x;
                    io.put(node, p, xo, x, wk + 6 * p + r);
                }
                wave_lds_order();
                near_load(k, q + LU_AHEAD, far[k]);
                far[k] = io.node(bpos(q + 2 * LU_AHEAD));
            }
        }
    }
}

// x = M⁻¹ b (ACCUM: x += M⁻¹ b) along each chain of an unfinished solve
template <bool ACCUM>
EBD void chain_lu_solve_b(ChainDev cd, const double* __restrict__ b, double* __restrict__ x,
                                                     const KrylovCtl* kc, int direct_only, const VB& vb) {
    __shared__ __align__(16) double ex[LU_GROUPS * 12];
    const Lane6 ln = lane6(nullptr);
    const int64_t ch = (int64_t)vb.x * LU_CHAINS + (ln.group >> 1);
    const bool live = ln.r < 6 && ch < cd.n_chains && krylov_open(kc[cd.inst[ch]], direct_only);
    const LuHalf h = lu_half(cd, ch, live, ln.group);
    if (wave_max_i32(live ? h.n : 0) == 0) return;
    const int32_t smax = wave_max_i32(h.steps);
    const int r = ln.r < 6 ? ln.r : 0;
    struct Io {
        const ChainDev* cd;
        const double* bv;
        double* xv;
        int64_t s0;
        int r;
        EBD int64_t node(int32_t p) const { return cd->node[s0 + p]; }
        EBD double b(int64_t nd, int32_t) const { return bv[6 * nd + r]; }
        EBD double old(int64_t nd, int32_t) const { return ACCUM ? xv[6 * nd + r] : 0.0; }
        EBD void put(int64_t nd, int32_t, double xo, double v, double*) const { xv[6 * nd + r] = ACCUM ? xo + v : v; }
    };
    const Io io{&cd, b, x, h.s0, r};
    // the top group's exchange at ex + 24 c, its bottom group's 12 doubles on
    double* e = ex + 12 * ln.group;
    chain_lu_solve_group(cd, h, smax, r, live, cd.yw + 6 * h.s0, e, e + 6, io);
}
#ifndef EB_BODIES_ONLY
template <bool ACCUM>
__global__ void __launch_bounds__(64) chain_lu_solve(ChainDev cd, const double* __restrict__ b, double* __restrict__ x,
                                                     const KrylovCtl* kc, int direct_only) { chain_lu_solve_b<ACCUM>(cd, b, x, kc, direct_only, EB_VB); }
#endif

// ---- closed chains (Sherman-Morrison-Woodbury) ----------------------------------------------
// A chain whose two ends are coupled (a ring) has the matrix A = T + U M Uᵀ: T the block-
// tridiagonal matrix the factorisation inverts, U the 12 unit columns of the first and last
// positions' dofs, M = [[0, A_fl], [A_lf, 0]] the two coupling blocks (entries touching a fixed
// dof zeroed, as in T). With Z = T⁻¹U and G = UᵀZ, A⁻¹ = T⁻¹ − Z H Uᵀ T⁻¹ with
// H = (I + M G)⁻¹ M, so the preconditioner stays exact on a ring: one chain solve y = T⁻¹ b,
// then y −= Z (H (Uᵀ y)).
struct ClosureDev {
    const int32_t *chain, *fl, *lf;   // per closure: chain, blocks (first, last) and (last, first)
    const uint64_t *mfl, *mlf;        // their fixed-dof masks
    const int64_t* zoff;              // per closure: offset of its Z (12 columns × 6 n, chain order)
    double* Z;
    double* H;                        // 144 per closure, row-major
    int32_t n;
};

EBD int32_t closure_unit(int j, int32_t n) { return j < 6 ? j : 6 * (n - 1) + (j - 6); }

// column j = blockIdx.y of Z for closure blockIdx.x: the chain solve of U's column j (in LDS:
// 12 doubles per position)
template <int NT>
EBD void closure_columns_b(ChainDev cd, ClosureDev cl, const uint8_t* __restrict__ active, const VB& vb) {
    extern __shared__ __align__(16) double cl_smem[];
    const int c = vb.x, j = vb.y;
    const int32_t ch = cl.chain[c];
    if (!active[cd.inst[ch]]) return;   // frozen instance: its factors were not refreshed
    const int64_t s0 = cd.ptr[ch];
    const int32_t n = cd.ptr[ch + 1] - (int32_t)s0;
    double* bw = cl_smem;
    double* yw = cl_smem + 6 * (int64_t)n;
    const int32_t unit = closure_unit(j, n);
    for (int32_t t = threadIdx.x; t < 6 * n; t += NT) bw[t] = t == unit ? 1.0 : 0.0;
    __syncthreads();
    chain_solve_body<NT>(cd, s0, n, bw, yw);
    double* z = cl.Z + cl.zoff[c] + (int64_t)j * 6 * n;
    for (int32_t t = threadIdx.x; t < 6 * n; t += NT) z[t] = yw[t];
}
#ifndef EB_BODIES_ONLY
template <int NT>
__global__ void __launch_bounds__(NT) closure_columns(ChainDev cd, ClosureDev cl, const uint8_t* __restrict__ active) { closure_columns_b<NT>(cd, cl, active, EB_VB); }
#endif

// the columns of Z with the block-LU factors: one chain's two six-lane groups per (closure,
// column), the column solved in Z
EBD void closure_columns_lu_b(ChainDev cd, ClosureDev cl, const uint8_t* __restrict__ active, const VB& vb) {
    __shared__ __align__(16) double ex[LU_GROUPS * 12];
    const Lane6 ln = lane6(nullptr);
    const int64_t item = (int64_t)vb.x * LU_CHAINS + (ln.group >> 1);
    const bool mine = ln.r < 6 && item < 12 * (int64_t)cl.n;
    const int c = mine ? (int)(item / 12) : 0, j = (int)(item % 12);
    const bool live = mine && active[cd.inst[cl.chain[c]]];
    const LuHalf h = lu_half(cd, live ? cl.chain[c] : 0, live, ln.group);
    if (wave_max_i32(live ? h.n : 0) == 0) return;
    const int32_t smax = wave_max_i32(h.steps);
    const int r = ln.r < 6 ? ln.r : 0;
    // lanes without a column read (never write) the start of Z
    double* z = live ? cl.Z + cl.zoff[c] + (int64_t)j * 6 * h.n : cl.Z;
    struct Io {
        int32_t unit;
        int r;
        EBD int64_t node(int32_t p) const { return p; }
        EBD double b(int64_t, int32_t p) const { return 6 * p + r == unit ? 1.0 : 0.0; }
        EBD double old(int64_t, int32_t) const { return 0.0; }
        EBD void put(int64_t, int32_t, double, double v, double* w) const { *w = v; }
    };
    const Io io{live ? closure_unit(j, h.n) : -1, r};
    double* e = ex + 12 * ln.group;
    chain_lu_solve_group(cd, h, smax, r, live, z, e, e + 6, io);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(64) closure_columns_lu(ChainDev cd, ClosureDev cl, const uint8_t* __restrict__ active) { closure_columns_lu_b(cd, cl, active, EB_VB); }
#endif

// H = (I + M G)⁻¹ M per closure, one wavefront each, in LDS: Gauss-Jordan elimination with
// partial pivoting on [I + M G | M]; a singular or non-finite system leaves H = 0 (no correction:
// the plain chain preconditioner)
EBD void closure_setup_b(ChainDev cd, ClosureDev cl, const double* __restrict__ A,
                                                    const uint8_t* __restrict__ active, const VB& vb) {
    __shared__ double M[12][12];
    __shared__ double W[12][24];
    __shared__ double fcol[12];
    __shared__ int piv_row, ok;
    // the work of one wavefront (called by a larger workgroup, the other lanes keep the barriers)
    const int c = vb.x, t = threadIdx.x, tw = t < 64 ? t : (1 << 20);
    const int32_t ch = cl.chain[c];
    if (!active[cd.inst[ch]]) return;
    const int32_t n = cd.ptr[ch + 1] - cd.ptr[ch];
    const double* Z = cl.Z + cl.zoff[c];
    for (int q = tw; q < 144; q += 64) {
        const int r = q / 12, k = q % 12;
        double v = 0.0;
        if (r < 6 && k >= 6) {
            const int e = 6 * r + (k - 6);
            v = ((cl.mfl[c] >> e) & 1) ? 0.0 : A[36 * (int64_t)cl.fl[c] + e];
        } else if (r >= 6 && k < 6) {
            const int e = 6 * (r - 6) + k;
            v = ((cl.mlf[c] >> e) & 1) ? 0.0 : A[36 * (int64_t)cl.lf[c] + e];
        }
        M[r][k] = v;
    }
    if (t == 0) ok = 1;
    __syncthreads();
    for (int q = tw; q < 144; q += 64) {
        const int r = q / 12, jc = q % 12;
        double v = r == jc ? 1.0 : 0.0;
        for (int k = 0; k < 12; ++k) v += M[r][k] * Z[(int64_t)jc * 6 * n + closure_unit(k, n)];
        W[r][jc] = v;
        W[r][12 + jc] = M[r][jc];
    }
    __syncthreads();
    for (int p = 0; p < 12; ++p) {
        if (t == 0) {
            int pr = p;
            for (int r = p + 1; r < 12; ++r)
                if (fabs(W[r][p]) > fabs(W[pr][p])) pr = r;
            piv_row = pr;
            if (!(fabs(W[pr][p]) > 0.0) || !isfinite(W[pr][p])) ok = 0;
        }
        __syncthreads();
        if (!ok) break;   // uniform: read after the barrier by every lane
        const int pr = piv_row;
        if (pr != p && t < 24) { const double a = W[p][t]; W[p][t] = W[pr][t]; W[pr][t] = a; }
        __syncthreads();
        const double inv = 1.0 / W[p][p];
        __syncthreads();
        if (t < 24) W[p][t] *= inv;
        if (t < 12) fcol[t] = W[t][p];   // the column's factors, read before any row changes
        __syncthreads();
        for (int q = tw; q < 12 * 24; q += 64) {
            const int r = q / 24, k = q % 24;
            if (r != p) W[r][k] -= fcol[r] * W[p][k];
        }
        __syncthreads();
    }
    double* H = cl.H + 144 * (int64_t)c;
    for (int q = tw; q < 144; q += 64) {
        const double v = W[q / 12][12 + q % 12];
        H[q] = ok && isfinite(v) ? v : 0.0;
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(64) closure_setup(ChainDev cd, ClosureDev cl, const double* __restrict__ A,
                                                    const uint8_t* __restrict__ active) { closure_setup_b(cd, cl, A, active, EB_VB); }
#endif

// y −= Z (H (Uᵀ y)) on each closed chain of an unfinished solve, after the chain solve
EBD void closure_apply_b(ChainDev cd, ClosureDev cl, double* __restrict__ y,
                                                     const KrylovCtl* __restrict__ kc, int direct_only, const VB& vb) {
    __shared__ double sv[12];
    const int c = vb.x;
    const int32_t ch = cl.chain[c];
    if (!krylov_open(kc[cd.inst[ch]], direct_only)) return;
    const int64_t s0 = cd.ptr[ch];
    const int32_t n = cd.ptr[ch + 1] - (int32_t)s0;
    if (threadIdx.x < 12) {
        const int32_t nf = cd.node[s0], nl = cd.node[s0 + n - 1];
        const double* H = cl.H + 144 * (int64_t)c + 12 * threadIdx.x;
        double acc = 0;
        for (int k = 0; k < 6; ++k) acc += H[k] * y[6 * (int64_t)nf + k];
        for (int k = 0; k < 6; ++k) acc += H[6 + k] * y[6 * (int64_t)nl + k];
        sv[threadIdx.x] = acc;
    }
    __syncthreads();
    const double* Z = cl.Z + cl.zoff[c];
    for (int32_t q = threadIdx.x; q < 6 * n; q += 256) {
        double acc = 0;
#pragma unroll
        for (int j = 0; j < 12; ++j) acc += Z[(int64_t)j * 6 * n + q] * sv[j];
        y[6 * (int64_t)cd.node[s0 + q / 6] + q % 6] -= acc;
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) closure_apply(ChainDev cd, ClosureDev cl, double* __restrict__ y,
                                                     const KrylovCtl* __restrict__ kc, int direct_only) { closure_apply_b(cd, cl, y, kc, direct_only, EB_VB); }
#endif

// ---- BiCGStab (van der Vorst), right-preconditioned, FP64 scalars on the device -------------
// One independent solve per instance, run in lockstep: the scalars live in KrylovCtl[instance],
// element-wise and reduction kernels work chunk by chunk, and the finalize kernels update every
// instance's scalars and the all-done flag the host polls.

// what the host polls: every instance done, the largest iteration count, and the first instance
// whose finished solve is not accepted (relative residual above 1e-6, or a breakdown short of
// the tolerance), with its relative residual
struct KrylovAll {
    int all_done, max_iters, fail_inst, breakdown;
    double fail_rel;
    int pivot_status;   // 1: a pivot block of some instance's factorisation is singular
    int n_failed;       // instances whose finished solve is not accepted
};

// start of a linear solve: clears the per-instance pivot flags the factorisation may set
EBD void solve_begin_b(int ninst, int32_t* pivot_inst, KrylovAll* all, const VB& vb) {
    for (int s = threadIdx.x; s < ninst; s += blockDim.x) pivot_inst[s] = 0;
    if (threadIdx.x == 0) all->pivot_status = 0;
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) solve_begin(int ninst, int32_t* pivot_inst, KrylovAll* all) { solve_begin_b(ninst, pivot_inst, all, EB_VB); }
#endif

// true when the kernel may run: ungated, or the polled solve is done
EBD bool gate_open(const KrylovAll* gate) { return !gate || gate->all_done; }

// acceptance of a finished solve: the tolerance reached, or a stall at rounding level (< 1e-6)
EBD bool krylov_failed(const KrylovCtl& k, double& rel) {
    rel = 0;
    if (k.pivot) { rel = 1; return true; }
    if (!k.done || k.iters == 0) return false;
    if (k.bnorm2 == 0.0 || k.rr2 <= k.tol2 * k.bnorm2 * 1.0000001 || k.ss2 <= k.tol2 * k.bnorm2) return false;
    rel = k.bnorm2 > 0 ? sqrt(fmin(k.rr2, k.ss2 > 0 ? k.ss2 : k.rr2) / k.bnorm2) : 0.0;
    return !(rel < 1e-6);
}

// folds every instance's state into the polled record (one FIN_BLOCK workgroup)
EBD void krylov_summary(const KrylovCtl* kc, int ninst, KrylovAll* all) {
    __shared__ int s_done, s_iters, s_fail, s_nfail, s_piv;
    if (threadIdx.x == 0) { s_done = 1; s_iters = 0; s_fail = 0x7fffffff; s_nfail = 0; s_piv = 0; }
    __syncthreads();
    for (int s = threadIdx.x; s < ninst; s += (int)blockDim.x) {
        const KrylovCtl& k = kc[s];
        double rel;
        if (!k.done) atomicAnd(&s_done, 0);
        atomicMax(&s_iters, k.iters);
        if (krylov_failed(k, rel)) { atomicMin(&s_fail, s); atomicAdd(&s_nfail, 1); }
        if (k.pivot) atomicOr(&s_piv, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        all->all_done = s_done;
        all->max_iters = s_iters;
        all->n_failed = s_nfail;
        all->pivot_status = s_piv;
        all->fail_inst = s_fail == 0x7fffffff ? -1 : s_fail;
        double rel = 0;
        if (s_fail != 0x7fffffff) krylov_failed(kc[s_fail], rel);
        all->fail_rel = rel;
        all->breakdown = s_fail != 0x7fffffff ? kc[s_fail].breakdown : 0;
    }
}

// x = 0, r = r̂ = b (b = residual with fixed dofs zeroed), ‖b‖²
EBD void krylov_init_b(Segs sg, const uint8_t* fixed, const double* res, double* x,
                                                         double* r, double* rhat, double* p, double* v, double* partials, const VB& vb) {
    double q[1] = {0};
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK) {
        double bi = fixed[i] ? 0.0 : res[i];
        x[i] = 0; r[i] = bi; rhat[i] = bi; p[i] = 0; v[i] = 0;
        q[0] += bi * bi;
    }
    seg_block_store<1>(q, partials, sg.nchunks, vb.x);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) krylov_init(Segs sg, const uint8_t* fixed, const double* res, double* x,
                                                         double* r, double* rhat, double* p, double* v, double* partials) { krylov_init_b(sg, fixed, res, x, r, rhat, p, v, partials, EB_VB); }
#endif
// instances that are not active in the Newton iteration start done
// instances whose factorisation met a singular pivot block start done and failed
EBD void krylov_init_fin_b(const double* partials, Segs sg, KrylovCtl* kc, double tol,
                                                             const uint8_t* active, const int32_t* pivot_inst,
                                                             const uint8_t* exact, KrylovAll* all, const VB& vb) {
    for (int s = threadIdx.x; s < sg.ninst; s += (int)blockDim.x) {
        double o[1];
        seg_sum<1>(partials, sg, s, o);
        KrylovCtl& k = kc[s];
        k.bnorm2 = o[0]; k.rho = o[0]; k.rho_old = 1; k.alpha = 1; k.omega = 1;
        k.rr2 = o[0]; k.ss2 = 0; k.tol2 = tol * tol; k.iters = 0; k.breakdown = 0;
        k.pivot = (active[s] && pivot_inst[s]) ? 1 : 0;
        k.direct = exact[s];
        k.done = (o[0] == 0.0 || !active[s] || k.pivot) ? 1 : 0;
    }
    __syncthreads();
    krylov_summary(kc, sg.ninst, all);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(FIN_BLOCK) krylov_init_fin(const double* partials, Segs sg, KrylovCtl* kc, double tol,
                                                             const uint8_t* active, const int32_t* pivot_inst,
                                                             const uint8_t* exact, KrylovAll* all) { krylov_init_fin_b(partials, sg, kc, tol, active, pivot_inst, exact, all, EB_VB); }
#endif
// p = r + β(p - ωv), β = (ρ/ρ_old)(α/ω); first iteration p = r
EBD void krylov_p_b(Segs sg, const KrylovCtl* kc, const double* r, double* p,
                                                      const double* v, const VB& vb) {
    const KrylovCtl& k = kc[sg.inst[vb.x]];
    if (k.done) return;
    const bool first = k.iters == 0;
    const double beta = first ? 0.0 : (k.rho / k.rho_old) * (k.alpha / k.omega);
    const double om = k.omega;
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK)
        p[i] = first ? r[i] : r[i] + beta * (p[i] - om * v[i]);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) krylov_p(Segs sg, const KrylovCtl* kc, const double* r, double* p,
                                                      const double* v) { krylov_p_b(sg, kc, r, p, v, EB_VB); }
#endif
EBD void krylov_alpha_fin_b(const double* partials, Segs sg, KrylovCtl* kc, const VB& vb) {
    for (int s = threadIdx.x; s < sg.ninst; s += (int)blockDim.x) {
        KrylovCtl& k = kc[s];
        if (k.done) continue;
        double o[1];
        seg_sum<1>(partials, sg, s, o);
        if (o[0] == 0.0 || !isfinite(o[0])) { k.breakdown = 1; k.done = 1; continue; }
        k.alpha = k.rho / o[0];
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(FIN_BLOCK) krylov_alpha_fin(const double* partials, Segs sg, KrylovCtl* kc) { krylov_alpha_fin_b(partials, sg, kc, EB_VB); }
#endif
// s = r - αv, x += αy, ‖s‖²
EBD void krylov_s_b(Segs sg, const KrylovCtl* kc, const double* r, const double* v,
                                                      const double* y, double* s, double* x, double* partials, const VB& vb) {
    const KrylovCtl& k = kc[sg.inst[vb.x]];
    if (k.done) return;
    const double al = k.alpha;
    double q[1] = {0};
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK) {
        double si = r[i] - al * v[i];
        s[i] = si;
        x[i] += al * y[i];
        q[0] += si * si;
    }
    seg_block_store<1>(q, partials, sg.nchunks, vb.x);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) krylov_s(Segs sg, const KrylovCtl* kc, const double* r, const double* v,
                                                      const double* y, double* s, double* x, double* partials) { krylov_s_b(sg, kc, r, v, y, s, x, partials, EB_VB); }
#endif
EBD void krylov_s_fin_b(const double* partials, Segs sg, KrylovCtl* kc, const VB& vb) {
    for (int s = threadIdx.x; s < sg.ninst; s += (int)blockDim.x) {
        KrylovCtl& k = kc[s];
        if (k.done) continue;
        double o[1];
        seg_sum<1>(partials, sg, s, o);
        k.ss2 = o[0];
        if (o[0] <= k.tol2 * k.bnorm2) { k.rr2 = o[0]; k.iters += 1; k.done = 1; }
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(FIN_BLOCK) krylov_s_fin(const double* partials, Segs sg, KrylovCtl* kc) { krylov_s_fin_b(partials, sg, kc, EB_VB); }
#endif
EBD void krylov_omega_fin_b(const double* partials, Segs sg, KrylovCtl* kc, const VB& vb) {
    for (int s = threadIdx.x; s < sg.ninst; s += (int)blockDim.x) {
        KrylovCtl& k = kc[s];
        if (k.done) continue;
        double o[2];
        seg_sum<2>(partials, sg, s, o);
        if (o[1] == 0.0 || !isfinite(o[1])) { k.breakdown = 1; k.done = 1; continue; }
        k.omega = o[0] / o[1];
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(FIN_BLOCK) krylov_omega_fin(const double* partials, Segs sg, KrylovCtl* kc) { krylov_omega_fin_b(partials, sg, kc, EB_VB); }
#endif
// x += ωz, r = s - ωt, (r̂·r), ‖r‖²
EBD void krylov_x_b(Segs sg, const KrylovCtl* kc, const double* z, const double* s,
                                                      const double* t, const double* rhat, double* x, double* r,
                                                      double* partials, const VB& vb) {
    const KrylovCtl& k = kc[sg.inst[vb.x]];
    if (k.done) return;
    const double om = k.omega;
    double q[2] = {0, 0};
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK) {
        x[i] += om * z[i];
        double ri = s[i] - om * t[i];
        r[i] = ri;
        q[0] += rhat[i] * ri;
        q[1] += ri * ri;
    }
    seg_block_store<2>(q, partials, sg.nchunks, vb.x);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) krylov_x(Segs sg, const KrylovCtl* kc, const double* z, const double* s,
                                                      const double* t, const double* rhat, double* x, double* r,
                                                      double* partials) { krylov_x_b(sg, kc, z, s, t, rhat, x, r, partials, EB_VB); }
#endif
EBD void krylov_rho_fin_b(const double* partials, Segs sg, KrylovCtl* kc, KrylovAll* all, const VB& vb) {
    for (int s = threadIdx.x; s < sg.ninst; s += (int)blockDim.x) {
        KrylovCtl& k = kc[s];
        if (!k.done) {
            double o[2];
            seg_sum<2>(partials, sg, s, o);
            k.rho_old = k.rho;
            k.rho = o[0];
            k.rr2 = o[1];
            k.iters += 1;
            if (o[1] <= k.tol2 * k.bnorm2) k.done = 1;
            else if (o[0] == 0.0 || !isfinite(o[0])) { k.breakdown = 1; k.done = 1; }
        }
    }
    __syncthreads();
    krylov_summary(kc, sg.ninst, all);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(FIN_BLOCK) krylov_rho_fin(const double* partials, Segs sg, KrylovCtl* kc, KrylovAll* all) { krylov_rho_fin_b(partials, sg, kc, all, EB_VB); }
#endif
// ---- the direct first step ----------------------------------------------------------------------
// When the chain cover (with the closure corrections) holds every coupling of the tangent, the
// preconditioner M is the solve's operator A up to rounding and x = M⁻¹ b solves the system. The
// solve of such an instance (KrylovCtl::direct, a property of the instance alone: no coupling left
// to a second chain cover) starts with that step: x = M⁻¹ b (precond_apply), r = b − A x, and ends
// where ‖r‖ meets the tolerance; if it does not, BiCGStab continues from that x (r₀ = r, r̂ = b, as
// after krylov_init). The other instances (nets) run BiCGStab from x = 0.
//
// r = b − A x on the chunks of unfinished direct solves (b = r̂, which krylov_init set to the
// right-hand side); partials: r̂·r, ‖r‖²
EBD void krylov_direct_residual_b(Segs sg, const int32_t* __restrict__ bptr,
                                                                    const int32_t* __restrict__ bcol,
                                                                    const uint8_t* __restrict__ fixed,
                                                                    const double* __restrict__ A, const double* __restrict__ x,
                                                                    const double* __restrict__ b, double* __restrict__ r,
                                                                    double* partials, const KrylovCtl* kc, const VB& vb) {
    if (!krylov_open(kc[sg.inst[vb.x]], 1)) return;
    double q[2] = {0, 0};
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK) {
        const double ri = b[i] - masked_row(i, bptr, bcol, fixed, A, x);
        r[i] = ri;
        q[0] += b[i] * ri;
        q[1] += ri * ri;
    }
    seg_block_store<2>(q, partials, sg.nchunks, vb.x);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) krylov_direct_residual(Segs sg, const int32_t* __restrict__ bptr,
                                                                    const int32_t* __restrict__ bcol,
                                                                    const uint8_t* __restrict__ fixed,
                                                                    const double* __restrict__ A, const double* __restrict__ x,
                                                                    const double* __restrict__ b, double* __restrict__ r,
                                                                    double* partials, const KrylovCtl* kc) { krylov_direct_residual_b(sg, bptr, bcol, fixed, A, x, b, r, partials, kc, EB_VB); }
#endif
EBD void krylov_direct_fin_b(const double* partials, Segs sg, KrylovCtl* kc,
                                                               KrylovAll* all, const VB& vb) {
    for (int s = threadIdx.x; s < sg.ninst; s += (int)blockDim.x) {
        KrylovCtl& k = kc[s];
        if (!krylov_open(k, 1)) continue;
        double o[2];
        seg_sum<2>(partials, sg, s, o);
        k.rr2 = o[1];
        if (o[1] <= k.tol2 * k.bnorm2) {
            k.iters = 1;
            k.done = 1;
        } else {
            k.rho = o[0];   // BiCGStab from this x: p = r at its first iteration (iters = 0)
        }
    }
    __syncthreads();
    krylov_summary(kc, sg.ninst, all);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(FIN_BLOCK) krylov_direct_fin(const double* partials, Segs sg, KrylovCtl* kc,
                                                               KrylovAll* all) { krylov_direct_fin_b(partials, sg, kc, all, EB_VB); }
#endif

// ΔD[free] = x[free] for the instances active in the Newton iteration
// Kernels of the corrector's update that take a `gate` run only once the linear solve is done: the
// corrector launches them behind a first Krylov iteration without waiting for the host, and
// launches them again after further iterations when that one did not finish the solve.
EBD void scatter_solution_b(int64_t ndofs, const uint8_t* fixed, const double* x, double* dD,
                                                        Active act, const KrylovAll* gate, const VB& vb) {
    if (!gate_open(gate)) return;
    int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (i < ndofs && !fixed[i] && act.flag[act.node_inst[i / 6]]) dD[i] = x[i];
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) scatter_solution(int64_t ndofs, const uint8_t* fixed, const double* x, double* dD,
                                                        Active act, const KrylovAll* gate) { scatter_solution_b(ndofs, fixed, x, dD, act, gate, EB_VB); }
#endif

// Instances whose solve failed (a singular pivot block, a breakdown or a stall short of the
// tolerance, or the iteration cap reached) leave the Newton iteration: the corrector records
// k = kfail (max_iterations + 1, a step that did not converge: Δt is halved and the step
// repeated, run_simulation.jl:55-71), the predictor flags them in pfail. The other instances
// carry on as if alone.
EBD void solve_screen_b(int ninst, const KrylovCtl* kc, uint8_t* active, int32_t* kiter,
                                                    int32_t kfail, uint8_t* pfail, const KrylovAll* gate, const VB& vb) {
    if (!gate_open(gate)) return;
    for (int s = threadIdx.x; s < ninst; s += blockDim.x) {
        if (!active[s]) continue;
        double rel;
        if (!kc[s].done || krylov_failed(kc[s], rel)) {
            active[s] = 0;
            if (pfail) pfail[s] = 1;
            else kiter[s] = kfail;
        }
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) solve_screen(int ninst, const KrylovCtl* kc, uint8_t* active, int32_t* kiter,
                                                    int32_t kfail, uint8_t* pfail, const KrylovAll* gate) { solve_screen_b(ninst, kc, active, kiter, kfail, pfail, gate, EB_VB); }
#endif

// ---- Newmark / rotation updates (src/utils_solver.jl:6-48, :157-245) -----------------------

struct NodeFields { double *u, *ud, *udd, *w, *wd, *wdd, *R, *dR, *u_n, *ud_n, *udd_n, *w_n, *wd_n, *wdd_n, *R_n, *dR_n; };

EBD void stv(double* a, int64_t n, V3<double> v) { a[3 * n] = v.x; a[3 * n + 1] = v.y; a[3 * n + 2] = v.z; }
EBD V3<double> ldv3(const double* a, int64_t n) { return {a[3 * n], a[3 * n + 1], a[3 * n + 2]}; }
EBD M3<double> ldm3(const double* a, int64_t n) { M3<double> m; for (int k = 0; k < 9; ++k) m.a[k] = a[9 * n + k]; return m; }
EBD void stm(double* a, int64_t n, const M3<double>& m) { for (int k = 0; k < 9; ++k) a[9 * n + k] = m.a[k]; }

// initialize_global_corrector! (predictor=0) / initialize_global_predictor! (predictor=1)
EBD void init_global_b(int64_t N, NodeFields nf, double* D, double* Dd, double* Ddd, double* dD,
                                                   int predictor, Active act, const VB& vb) {
    int64_t n = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (n >= N || !act.flag[act.node_inst[n]]) return;
    for (int k = 0; k < 3; ++k) {
        dD[6 * n + k] = 0; D[6 * n + k] = nf.u[3 * n + k]; Dd[6 * n + k] = nf.ud[3 * n + k]; Ddd[6 * n + k] = nf.udd[3 * n + k];
        dD[6 * n + 3 + k] = 0; D[6 * n + 3 + k] = predictor ? 0.0 : nf.w[3 * n + k];
        Dd[6 * n + 3 + k] = nf.wd[3 * n + k]; Ddd[6 * n + 3 + k] = nf.wdd[3 * n + k];
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) init_global(int64_t N, NodeFields nf, double* D, double* Dd, double* Ddd, double* dD,
                                                   int predictor, Active act) { init_global_b(N, nf, D, Dd, Ddd, dD, predictor, act, EB_VB); }
#endif

// compute_global_corrector! per dof (all dofs, fixed included)
EBD void global_corrector_b(int64_t ndofs, const StepCoef* __restrict__ coef, double* D,
                                                        double* Dd, double* Ddd, const double* dD, Active act,
                                                        const KrylovAll* gate, const VB& vb) {
    if (!gate_open(gate)) return;
    int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    const int32_t s = act.node_inst[i / 6];
    if (!act.flag[s]) return;
    const double gb = coef[s].gbdt, inv = coef[s].inv;
    D[i] += dD[i]; Dd[i] += gb * dD[i]; Ddd[i] += inv * dD[i];
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) global_corrector(int64_t ndofs, const StepCoef* __restrict__ coef, double* D,
                                                        double* Dd, double* Ddd, const double* dD, Active act,
                                                        const KrylovAll* gate) { global_corrector_b(ndofs, coef, D, Dd, Ddd, dD, act, gate, EB_VB); }
#endif
// compute_global_predictor!
EBD void global_predictor_b(int64_t ndofs, const StepCoef* __restrict__ coef, double* D,
                                                        double* Dd, double* Ddd, const double* dD, Active act, const VB& vb) {
    int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    const int32_t s = act.node_inst[i / 6];
    if (!act.flag[s]) return;
    const StepCoef& sc = coef[s];
    const double dt = sc.dt, gob = sc.gob, gbdt = sc.gbdt, inv = sc.inv, dt2b = sc.dt2b;
    double Ddn = Dd[i];
    D[i] += dD[i];
    Dd[i] += (gbdt * dD[i] - gob * Ddn) + dt2b * Ddd[i];
    Ddd[i] += inv * ((dD[i] - dt * Ddn) - ((dt * dt) / 2) * Ddd[i]);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) global_predictor(int64_t ndofs, const StepCoef* __restrict__ coef, double* D,
                                                        double* Dd, double* Ddd, const double* dD, Active act) { global_predictor_b(ndofs, coef, D, Dd, Ddd, dD, act, EB_VB); }
#endif

// compute_local_corrector!: u, u̇, ü from D; ΔR ← exp(ΔD_θ)ΔR; w = log ΔR; ẇ, ẅ; R = ΔR Rⁿ
EBD void local_corrector_b(int64_t N, NodeFields nf, const double* D, const double* Dd,
                                                       const double* Ddd, const double* dD,
                                                       const StepCoef* __restrict__ coef, Active act,
                                                       const KrylovAll* gate, const VB& vb) {
    if (!gate_open(gate)) return;
    int64_t n = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (n >= N || !act.flag[act.node_inst[n]]) return;
    const HHTc h = coef[act.node_inst[n]].h;
    stv(nf.u, n, {D[6 * n], D[6 * n + 1], D[6 * n + 2]});
    stv(nf.ud, n, {Dd[6 * n], Dd[6 * n + 1], Dd[6 * n + 2]});
    stv(nf.udd, n, {Ddd[6 * n], Ddd[6 * n + 1], Ddd[6 * n + 2]});
    M3<double> dR = rotation_matrix({dD[6 * n + 3], dD[6 * n + 4], dD[6 * n + 5]}) * ldm3(nf.dR, n);
    stm(nf.dR, n, dR);
    V3<double> wdn = ldv3(nf.wd_n, n), wddn = ldv3(nf.wdd_n, n);
    V3<double> w = toangle(dR);
    V3<double> a = ((h.c_w * w + h.c_wd * wdn) + h.c_wdd * wddn);
    V3<double> c = ((h.d_w * w - h.d_wd * wdn) - h.d_wdd * wddn);
    stv(nf.wd, n, dR * a);
    stv(nf.wdd, n, dR * c);
    stm(nf.R, n, dR * ldm3(nf.R_n, n));
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) local_corrector(int64_t N, NodeFields nf, const double* D, const double* Dd,
                                                       const double* Ddd, const double* dD,
                                                       const StepCoef* __restrict__ coef, Active act,
                                                       const KrylovAll* gate) { local_corrector_b(N, nf, D, Dd, Ddd, dD, coef, act, gate, EB_VB); }
#endif

// compute_local_predictor!: w = θ̃ = D_θ, ΔR = exp(θ̃), R = ΔR Rⁿ, ẇ, ẅ from θ̃
EBD void local_predictor_b(int64_t N, NodeFields nf, const double* D, const double* Dd,
                                                       const double* Ddd, const StepCoef* __restrict__ coef, Active act, const VB& vb) {
    int64_t n = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (n >= N || !act.flag[act.node_inst[n]]) return;
    const HHTc h = coef[act.node_inst[n]].h;
    stv(nf.u, n, {D[6 * n], D[6 * n + 1], D[6 * n + 2]});
    stv(nf.ud, n, {Dd[6 * n], Dd[6 * n + 1], Dd[6 * n + 2]});
    stv(nf.udd, n, {Ddd[6 * n], Ddd[6 * n + 1], Ddd[6 * n + 2]});
    V3<double> th = {D[6 * n + 3], D[6 * n + 4], D[6 * n + 5]};
    stv(nf.w, n, th);
    M3<double> dR = rotation_matrix(th);
    stm(nf.dR, n, dR);
    stm(nf.R, n, dR * ldm3(nf.R_n, n));
    V3<double> wdn = ldv3(nf.wd_n, n), wddn = ldv3(nf.wdd_n, n);
    V3<double> a = ((h.c_w * th + h.c_wd * wdn) + h.c_wdd * wddn);
    V3<double> c = ((h.d_w * th - h.d_wd * wdn) - h.d_wdd * wddn);
    stv(nf.wd, n, dR * a);
    stv(nf.wdd, n, dR * c);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) local_predictor(int64_t N, NodeFields nf, const double* D, const double* Dd,
                                                       const double* Ddd, const StepCoef* __restrict__ coef, Active act) { local_predictor_b(N, nf, D, Dd, Ddd, coef, act, EB_VB); }
#endif

// compute_norms_corrector (utils_solver.jl:126-151): Σ over free dofs of ΔD², r², (f+Tᶜ+Tᵏ)² and
// Σ over fixed dofs of (f+Tᶜ+Tᵏ)², per instance
EBD void norms_partials_b(Segs sg, const uint8_t* fixed, const double* dD,
                                                            const double* res, const double* fext, const double* Tc,
                                                            const double* Tk, double* partials, const VB& vb) {
    double q[4] = {0, 0, 0, 0};
    for (int64_t i = sg.lo[vb.x] + threadIdx.x; i < sg.hi[vb.x]; i += RED_BLOCK) {
        double f = (fext[i] + Tc[i]) + Tk[i];
        if (fixed[i]) {
            q[3] += f * f;
        } else {
            q[0] += dD[i] * dD[i];
            q[1] += res[i] * res[i];
            q[2] += f * f;
        }
    }
    seg_block_store<4>(q, partials, sg.nchunks, vb.x);
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(RED_BLOCK) norms_partials(Segs sg, const uint8_t* fixed, const double* dD,
                                                            const double* res, const double* fext, const double* Tc,
                                                            const double* Tk, double* partials) { norms_partials_b(sg, fixed, dD, res, fext, Tc, Tk, partials, EB_VB); }
#endif

struct NewtonCtl {            // corrector loop control (src/corrector.jl:17-36)
    double tol_res, tol_disp;
    int k, max_iterations;    // k: set on the device from the count of finalised iterations
    int newton;               // 0: norms only (no convergence bookkeeping)
};
struct NewtonAll { int n_active, pad[3]; };

// per-instance norms (relative residual, ‖ΔD‖) into out[2s], out[2s+1]; in a Newton iteration,
// an active instance whose norms pass (or that has used its iterations) is frozen and its k
// recorded as the reference's loop leaves it (iterations done + 1)
// In a Newton iteration the iteration number is kept on the device: `kdone` counts the
// iterations finalised since newton_reset, so that one launch sequence serves every iteration.
EBD void norms_finalize_b(const double* partials, Segs sg, double* out, NewtonCtl nc,
                                                            uint8_t* active, int32_t* kiter, NewtonAll* all,
                                                            const KrylovAll* gate, int32_t* kdone, const VB& vb) {
    if (!gate_open(gate)) return;
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    if (nc.newton) nc.k = *kdone + 1;
    __syncthreads();
    int mine = 0;
    for (int s = threadIdx.x; s < sg.ninst; s += (int)blockDim.x) {
        if (nc.newton && !active[s]) continue;
        double o[4];
        seg_sum<4>(partials, sg, s, o);
        double dn = sqrt(o[0]), rn = sqrt(o[1]), fn = sqrt(o[2]), en = sqrt(o[3]);
        if (fn + en > 1e-12) rn = rn / (fn + en);
        out[2 * s] = rn;
        out[2 * s + 1] = dn;
        if (nc.newton) {
            const bool more = (rn > nc.tol_res || dn > nc.tol_disp) && nc.k + 1 <= nc.max_iterations;
            if (more) {
                mine += 1;
            } else {
                active[s] = 0;
                kiter[s] = nc.k + 1;
            }
        }
    }
    if (mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        all->n_active = cnt;
        if (nc.newton) *kdone = nc.k;
    }
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(FIN_BLOCK) norms_finalize(const double* partials, Segs sg, double* out, NewtonCtl nc,
                                                            uint8_t* active, int32_t* kiter, NewtonAll* all,
                                                            const KrylovAll* gate, int32_t* kdone) { norms_finalize_b(partials, sg, out, nc, active, kiter, all, gate, kdone, EB_VB); }
#endif

// Start (and end) of a corrector loop: the instances taking part in the step become active. At
// the start (`pfail` given) an instance whose predictor solve failed stays frozen with k = kfail.
EBD void newton_reset_b(int ninst, const uint8_t* part, const uint8_t* pfail, uint8_t* active, int32_t* kiter,
                             int32_t* kdone, int32_t kfail, const VB& vb) {
    if (vb.x == 0 && threadIdx.x == 0) *kdone = 0;
    for (int s = vb.x * blockDim.x + threadIdx.x; s < ninst; s += vb.n * blockDim.x) {
        const bool failed = pfail && part[s] && pfail[s];
        active[s] = (part[s] && !failed) ? 1 : 0;
        kiter[s] = failed ? kfail : 0;
    }
}
#ifndef EB_BODIES_ONLY
__global__ void newton_reset(int ninst, const uint8_t* part, const uint8_t* pfail, uint8_t* active, int32_t* kiter,
                             int32_t* kdone, int32_t kfail) { newton_reset_b(ninst, part, pfail, active, kiter, kdone, kfail, EB_VB); }
#endif

// several device copies in one launch (the commit and the rollback of a step): list y = blockIdx.y.
// With `sel`, element i of list y is copied only when its instance, map[y][i / per[y]], is selected.
struct CopyList {
    const double* src[16];
    double* dst[16];
    int64_t n[16];
    const int32_t* map[16];
    int32_t per[16];
    const uint8_t* sel;
};
EBD void copy_many_b(CopyList cl, const VB& vb) {
    const int y = vb.y;
    for (int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x; i < cl.n[y]; i += (int64_t)vb.n * blockDim.x)
        if (!cl.sel || cl.sel[cl.map[y][i / cl.per[y]]]) cl.dst[y][i] = cl.src[y][i];
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) copy_many(CopyList cl) { copy_many_b(cl, EB_VB); }
#endif

// ---- layout conversions at the ABI ----------------------------------------------------------
EBD void gather_idx32_b(int64_t n, const int32_t* idx, const double* src, double* dst, const VB& vb) {
    int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) gather_idx32(int64_t n, const int32_t* idx, const double* src, double* dst) { gather_idx32_b(n, idx, src, dst, EB_VB); }
#endif
EBD void scatter_idx32_b(int64_t n, const int32_t* idx, const double* src, double* dst, const VB& vb) {
    int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}
#ifndef EB_BODIES_ONLY
__global__ void __launch_bounds__(256) scatter_idx32(int64_t n, const int32_t* idx, const double* src, double* dst) { scatter_idx32_b(n, idx, src, dst, EB_VB); }
#endif

}  // namespace ebd
