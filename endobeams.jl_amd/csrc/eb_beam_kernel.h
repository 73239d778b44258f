// eb_beam_kernel.h — corotational beam element (compute_beams, src/beams/compute_beams.jl:1-538)
// as a gfx950 kernel: one lane per element. Element constants are component-major ([k][E]) so a
// wavefront's loads are contiguous rows; outputs are element-major, each 12x12 matrix as its four
// 6x6 node blocks, row-major (stage_idx), so that the assembly gather reads a block's 36 entries
// as consecutive values.
//
// Precision: the local frame, R̅ = Rₑᵀ R Rₑ⁰, the log map Θ̅, Tₛ⁻¹(Θ̅), η/μ and K̄ₕ are always FP64;
// everything downstream of them runs in the physics type P (double for the parity form, float
// for the precision form) and is stored as P, except the forces Tⁱⁿᵗ and Tᵏ, which enter the
// residual and stay FP64.
#pragma once
#include "eb_dmath.h"

namespace ebd {

struct NodeState {  // device pointers, Julia layouts (3N node-major vectors, 9N column-major matrices)
    const double *X0, *u, *ud, *udd, *wd, *wdd, *R, *dR;
};

struct BeamConsts {  // component-major element constants
    const int32_t *n1, *n2;  // 0-based
    const double* l0;        // [E]
    const double* Re0;       // [9][E]
    const double* props;     // [14][E]
    int64_t E;
};

template <typename S> struct BeamStage {  // element-major outputs: beam e owns a contiguous record
    S *K, *C, *M;        // [E][144] in stage_idx order (local dof order u₁ θ₁ u₂ θ₂), storage type S
    double *Tint, *Tk;   // [E][12]: forces are always FP64 (they make the residual)
    double* energy;      // [E][3]: strain, kinetic, contact
};

struct GaussPts { int n; double w[8], z[8]; };

EBD V3<double> ldv(const double* a, int64_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
EBD M3<double> ldm(const double* a, int64_t i) { M3<double> m; for (int k = 0; k < 9; ++k) m.a[k] = a[9 * i + k]; return m; }

// entry (R, C) of beam e's 12x12 matrix in the staging: node block (R/6, C/6) at position
// 2(R/6) + C/6 of the beam's record, each block row-major
EBD int64_t stage_idx(int64_t e, int R, int C) { return e * 144 + (2 * (R / 6) + C / 6) * 36 + (R % 6) * 6 + C % 6; }

// The element kernel writes a tile's K and M records (64 beams × 144 entries) to LDS first, one row of
// STAGE_REC entries per beam (an odd stride: the lanes' scattered writes hit distinct banks), and
// copies them out as 16-byte pieces: the tile's records are contiguous in the staging, so the copy
// is one coalesced stream instead of 144 stores per lane to 64 different cache lines.
constexpr int STAGE_REC = 145;
template <typename S> struct StPiece { S v[16 / sizeof(S)]; };
template <typename S, typename P> EBD void st_rec(S* rec, int bi, int bj, const M3<P>& B) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int R = 3 * bi + i, C = 3 * bj + j;
            rec[(2 * (R / 6) + C / 6) * 36 + (R % 6) * 6 + C % 6] = S(B.a[i + 3 * j]);
        }
}
// the wavefront's LDS writes in issue order before its lanes read each other's (the LDS executes a
// wavefront's instructions in order; this keeps the compiler from moving them)
EBD void wave_lds_fence() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}
// a wavefront copies the tile's records from LDS to the staging at dst (the tile's first record);
// vmask: bit j set when beam j of the tile was computed
template <typename S> EBD void tile_records_out(S* __restrict__ dst, const S* rec, uint64_t vmask, int lane) {
    constexpr int NP = 16 / sizeof(S), PR = 144 / NP;
    wave_lds_fence();
#pragma unroll 4
    for (int q = lane; q < 64 * PR; q += 64) {
        const int j = q / PR, k = q - PR * j;
        if ((vmask >> j) & 1) {
            StPiece<S> w;
#pragma unroll
            for (int c = 0; c < NP; ++c) w.v[c] = rec[j * STAGE_REC + NP * k + c];
            *reinterpret_cast<StPiece<S>*>(dst + (int64_t)j * 144 + NP * k) = w;
        }
    }
}

template <typename S, typename P> EBD void st_block(S* base, int64_t E, int64_t e, int bi, int bj, const M3<P>& B) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) base[stage_idx(e, 3 * bi + i, 3 * bj + j)] = S(B.a[i + 3 * j]);
}
template <typename P> EBD void st_vec(double* base, int64_t E, int64_t e, int b, V3<P> v) {
    base[e * 12 + 3 * b] = double(v.x); base[e * 12 + 3 * b + 1] = double(v.y); base[e * 12 + 3 * b + 2] = double(v.z);
}

struct ShapeF { double N1, N2, N3, N4, N5, N6, N7, N8; };
EBD ShapeF shape_fn(double l0, double z) {
    ShapeF s;
    double xi = l0 * (z + 1) / 2;
    s.N1 = 1 - xi / l0;
    s.N2 = 1 - s.N1;
    double a = 1 - xi / l0;
    s.N3 = xi * (a * a);
    s.N4 = -(1 - xi / l0) * ((xi * xi) / l0);
    s.N5 = (1 - 3 * xi / l0) * (1 - xi / l0);
    s.N6 = (3 * xi / l0 - 2) * (xi / l0);
    s.N7 = s.N3 + s.N4;
    s.N8 = s.N5 + s.N6 - 1;
    return s;
}

// Static (internal force / stiffness) part, compute_beams.jl:13-159. Writes Tⁱⁿᵗ and Kⁱⁿᵗ.
// Returns the quantities the dynamic part needs.
struct StaticOut {
    M3<double> Re;
    V3<double> Tb1, Tb2, v1;
    double eta, eta11, eta12, eta21, eta22, ln;
};

// no work beside the element (the element kernel without contact). A hook tests the Gauss points
// `gsel` of a beam with its frame (test) and, once every part of the tile is done, acts on the
// combined result (finish).
struct NoStaticHook {
    static constexpr bool SPLIT = false;
    EBD unsigned test(int64_t, const BeamConsts&, const GaussPts&, unsigned, const V3<double>&, const V3<double>&,
                      const M3<double>&, const V3<double>&, const V3<double>&) const { return 0; }
    EBD void finish(int64_t, unsigned) const {}
};

// The static part's tangent Kⁱⁿᵗ = BᵀK̄B + Kₘ + K̃ (compute_beams.jl:76-159) from the frame, Tₛ⁻¹,
// T̄ⁱⁿᵗ, K̄ₕ and B, all given in the physics type P; stored in the storage type of the staging.
template <typename P, typename S>
EBD void beam_static_tangent(S* krec, const M3<P>& Re, const V3<P>& v1, P eta, P ln,
                             P Tu, P Ku, const M3<P>& Ti1, const M3<P>& Ti2, const V3<P>& TT1, const V3<P>& TT2,
                             const M3<P>& Kh1, const M3<P>& Kh2, const M3<P>& G1, const M3<P>& G2, const M3<P>& G4,
                             const V3<P>& B1, const M3<P>& B11, const M3<P>& B12, const M3<P>& B14, const M3<P>& B21,
                             const M3<P>& B22, const M3<P>& B24, P KT0, P KT1, P KT2, P KTT0, P KTT1, P KTT2) {
    const M3<P> I = eye<P>();
    const M3<P> ReT = tr(Re);
    M3<P> P11 = -G1, P12 = I - G2, P22 = -G2, P14 = -G4, P24 = I - G4;
    M3<P> Bp11 = P11 * ReT, Bp12 = P12 * ReT, Bp22 = P22 * ReT, Bp14 = P14 * ReT, Bp24 = P24 * ReT;
    const M3<P>& Bp21 = Bp11;
    V3<P> Mp1 = tmul(Ti1, TT1), Mp2 = tmul(Ti2, TT2);
    M3<P> Q1 = skew(tmul(P11, Mp1) + tmul(P11, Mp2));
    M3<P> Q2 = skew(tmul(P12, Mp1) + tmul(P22, Mp2));
    M3<P> Q4 = skew(tmul(P14, Mp1) + tmul(P24, Mp2));
    V3<P> av = {P(0), eta * (Mp1.x + Mp2.x) / ln + (Mp1.y + Mp2.y) / ln, (Mp1.z + Mp2.z) / ln};
    M3<P> D3 = (I - outer(v1, v1)) / ln;
    V3<P> EGa1 = Re * tmul(G1, av);
    M3<P> Km11 = (D3 * Tu - (Re * (Q1 * G1)) * ReT) + outer(EGa1, B1);
    M3<P> Km12 = -((Re * (Q1 * G2)) * ReT);
    M3<P> Km14 = -((Re * (Q1 * G4)) * ReT);
    M3<P> Km22 = -((Re * (Q2 * G2)) * ReT);
    M3<P> Km24 = -((Re * (Q2 * G4)) * ReT);
    M3<P> Km44 = -((Re * (Q4 * G4)) * ReT);
    M3<P> tB11 = tr(Bp11), tB12 = tr(Bp12), tB14 = tr(Bp14), tB22 = tr(Bp22), tB24 = tr(Bp24);
    M3<P> Kt11 = ((tB11 * Kh1) * Bp11 + (tB11 * Kh2) * Bp21) + Km11;
    M3<P> Kt12 = ((tB11 * Kh1) * Bp12 + (tB11 * Kh2) * Bp22) + Km12;
    M3<P> Kt14 = ((tB11 * Kh1) * Bp14 + (tB11 * Kh2) * Bp24) + Km14;
    M3<P> Kt22 = ((tB12 * Kh1) * Bp12 + (tB22 * Kh2) * Bp22) + Km22;
    M3<P> Kt24 = ((tB12 * Kh1) * Bp14 + (tB22 * Kh2) * Bp24) + Km24;
    M3<P> Kt44 = ((tB14 * Kh1) * Bp14 + (tB24 * Kh2) * Bp24) + Km44;
    M3<P> KT = diag(KT0, KT1, KT2), KTT = diag(KTT0, KTT1, KTT2);
    M3<P> XB12 = KT * B12 + KTT * B22, YB12 = KTT * B12 + KT * B22;
    M3<P> XB14 = KT * B14 + KTT * B24, YB14 = KTT * B14 + KT * B24;
    M3<P> tb11 = tr(B11), tb21 = tr(B21), tb12 = tr(B12), tb22 = tr(B22), tb14 = tr(B14), tb24 = tr(B24);
    M3<P> K11 = ((outer(B1 * Ku, B1) + tb11 * (KT * B11 + KTT * B21)) + tb21 * (KTT * B11 + KT * B21)) + Kt11;
    M3<P> K12 = (tb11 * XB12 + tb21 * YB12) + Kt12;
    M3<P> K14 = (tb11 * XB14 + tb21 * YB14) + Kt14;
    M3<P> K22 = (tb12 * XB12 + tb22 * YB12) + Kt22;
    M3<P> K24 = (tb12 * XB14 + tb22 * YB14) + Kt24;
    M3<P> K44 = (tb14 * XB14 + tb24 * YB14) + Kt44;
    // block layout of Kⁱⁿᵗ (compute_beams.jl:148-159)
    M3<P> tK12 = tr(K12), tK14 = tr(K14);
    st_rec<S>(krec, 0, 0, cvt<S>(K11));
    st_rec<S>(krec, 1, 0, cvt<S>(tK12));
    st_rec<S>(krec, 2, 0, cvt<S>(-tr(K11)));      // K¹³ᵀ with K¹³ = -K¹¹
    st_rec<S>(krec, 3, 0, cvt<S>(tK14));
    st_rec<S>(krec, 0, 1, cvt<S>(K12));
    st_rec<S>(krec, 1, 1, cvt<S>(K22));
    st_rec<S>(krec, 2, 1, cvt<S>(-K12));          // K²³ᵀ = (-K¹²ᵀ)ᵀ
    st_rec<S>(krec, 3, 1, cvt<S>(tr(K24)));
    st_rec<S>(krec, 0, 2, cvt<S>(-K11));
    st_rec<S>(krec, 1, 2, cvt<S>(-tK12));
    st_rec<S>(krec, 2, 2, cvt<S>(K11));
    st_rec<S>(krec, 3, 2, cvt<S>(-tK14));         // K³⁴ᵀ = (-K¹⁴)ᵀ
    st_rec<S>(krec, 0, 3, cvt<S>(K14));
    st_rec<S>(krec, 1, 3, cvt<S>(K24));
    st_rec<S>(krec, 2, 3, cvt<S>(-K14));
    st_rec<S>(krec, 3, 3, cvt<S>(K44));
}

// krec: the beam's K record (LDS); with_frame(x₁, x₂, Rₑ, Θ̅₁, Θ̅₂) runs as soon as the frame is known
template <typename P, typename F>
EBD StaticOut beam_static(int64_t e, const NodeState& ns, const BeamConsts& bc, const BeamStage<P>& st, P* krec,
                          const F& with_frame) {
    const int64_t E = bc.E;
    const int32_t a = bc.n1[e], b = bc.n2[e];
    V3<double> x1 = ldv(ns.X0, a) + ldv(ns.u, a), x2 = ldv(ns.X0, b) + ldv(ns.u, b);
    double ln = norm(x2 - x1);
    double l0 = bc.l0[e];
    M3<double> Re0;
    for (int k = 0; k < 9; ++k) Re0.a[k] = bc.Re0[k * E + e];
    M3<double> R1 = ldm(ns.R, a), R2 = ldm(ns.R, b);
    LocalFrame f = local_frame(x1, x2, R1, R2, col(Re0, 1), ln);
    M3<double> ReT = tr(f.Re);
    M3<double> Rb1 = (ReT * R1) * Re0, Rb2 = (ReT * R2) * Re0;
    V3<double> Tb1 = toangle(Rb1), Tb2 = toangle(Rb2);
    with_frame(x1, x2, f.Re, Tb1, Tb2);
    M3<double> Ti1 = Ts_inv(Tb1), Ti2 = Ts_inv(Tb2);
    double pr[14];
    for (int k = 0; k < 14; ++k) pr[k] = bc.props[k * E + e];
    const double Ku = pr[2];
    const double KT0 = pr[3], KT1 = pr[4], KT2 = pr[5], KTT0 = pr[6], KTT1 = pr[7], KTT2 = pr[8];
    double ub = ln - l0;
    double Tu = Ku * ub;
    V3<double> TT1 = {KT0 * Tb1.x + KTT0 * Tb2.x, KT1 * Tb1.y + KTT1 * Tb2.y, KT2 * Tb1.z + KTT2 * Tb2.z};
    V3<double> TT2 = {KTT0 * Tb1.x + KT0 * Tb2.x, KTT1 * Tb1.y + KT1 * Tb2.y, KTT2 * Tb1.z + KT2 * Tb2.z};
    st.energy[3 * e] = (ub * Tu + dot(Tb1, TT1) + dot(Tb2, TT2)) / 2;
    double eta1, mu1, eta2, mu2;
    eta_mu(Tb1, eta1, mu1);
    eta_mu(Tb2, eta2, mu2);
    M3<double> Kh1 = Kh_bar(Tb1, TT1, Ti1, eta1, mu1);
    M3<double> Kh2 = Kh_bar(Tb2, TT2, Ti2, eta2, mu2);

    // Gᵀ matrices (local_Rₑ_and_aux): G1 = Gᵀu₁, G2 = GᵀΘ₁, G3 = -G1, G4 = GᵀΘ₂
    const M3<double> I = eye<double>();
    M3<double> G1 = rows(0.0, 0.0, f.eta / ln, 0.0, 0.0, 1 / ln, 0.0, -1 / ln, 0.0);
    M3<double> G2 = rows(f.eta12 / 2, -f.eta11 / 2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
    M3<double> G4 = rows(f.eta22 / 2, -f.eta21 / 2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
    M3<double> P11 = -G1, P12 = I - G2, P22 = -G2, P14 = -G4, P24 = I - G4;
    M3<double> Bp11 = P11 * ReT, Bp12 = P12 * ReT, Bp22 = P22 * ReT, Bp14 = P14 * ReT, Bp24 = P24 * ReT;
    const M3<double>& Bp21 = Bp11;
    V3<double> B1 = -f.v1;
    M3<double> B11 = Ti1 * Bp11, B12 = Ti1 * Bp12, B14 = Ti1 * Bp14;
    M3<double> B21 = Ti2 * Bp21, B22 = Ti2 * Bp22, B24 = Ti2 * Bp24;

    // Tⁱⁿᵗ = Bᵀ T̄ⁱⁿᵗ
    V3<double> T1 = (B1 * Tu + tmul(B11, TT1)) + tmul(B21, TT2);
    V3<double> T2 = tmul(B12, TT1) + tmul(B22, TT2);
    V3<double> T4 = tmul(B14, TT1) + tmul(B24, TT2);
    st_vec(st.Tint, E, e, 0, T1);
    st_vec(st.Tint, E, e, 1, T2);
    st_vec(st.Tint, E, e, 2, -T1);
    st_vec(st.Tint, E, e, 3, T4);

    // Kₘ, K̃, BᵀKB downstream of K̄ₕ: in the physics type P (FP64 in the parity form, the same
    // expressions in FP32 in the precision form)
    beam_static_tangent<P>(krec, cvt<P>(f.Re), cvt<P>(f.v1), P(f.eta), P(ln), P(Tu), P(Ku), cvt<P>(Ti1), cvt<P>(Ti2),
                           cvt<P>(TT1), cvt<P>(TT2), cvt<P>(Kh1), cvt<P>(Kh2), cvt<P>(G1), cvt<P>(G2), cvt<P>(G4),
                           cvt<P>(B1), cvt<P>(B11), cvt<P>(B12), cvt<P>(B14), cvt<P>(B21), cvt<P>(B22), cvt<P>(B24),
                           P(KT0), P(KT1), P(KT2), P(KTT0), P(KTT1), P(KTT2));

    StaticOut so;
    so.Re = f.Re; so.Tb1 = Tb1; so.Tb2 = Tb2; so.v1 = f.v1;
    so.eta = f.eta; so.eta11 = f.eta11; so.eta12 = f.eta12; so.eta21 = f.eta21; so.eta22 = f.eta22; so.ln = ln;
    return so;
}

// The element's corotational frame alone (the first lines of beam_static), for the parts of the
// element computed in other wavefronts.
EBD StaticOut beam_frame(int64_t e, const NodeState& ns, const BeamConsts& bc) {
    const int64_t E = bc.E;
    const int32_t a = bc.n1[e], b = bc.n2[e];
    V3<double> x1 = ldv(ns.X0, a) + ldv(ns.u, a), x2 = ldv(ns.X0, b) + ldv(ns.u, b);
    double ln = norm(x2 - x1);
    M3<double> Re0;
    for (int k = 0; k < 9; ++k) Re0.a[k] = bc.Re0[k * E + e];
    M3<double> R1 = ldm(ns.R, a), R2 = ldm(ns.R, b);
    LocalFrame f = local_frame(x1, x2, R1, R2, col(Re0, 1), ln);
    M3<double> ReT = tr(f.Re);
    StaticOut so;
    so.Re = f.Re;
    so.Tb1 = toangle((ReT * R1) * Re0);
    so.Tb2 = toangle((ReT * R2) * Re0);
    so.v1 = f.v1;
    so.eta = f.eta; so.eta11 = f.eta11; so.eta12 = f.eta12; so.eta21 = f.eta21; so.eta22 = f.eta22; so.ln = ln;
    return so;
}

}  // namespace ebd

namespace ebd {

// Parts of the dynamic computation, each run by its own wavefront of a workgroup
// (PART_MASS: Tᵏ, kinetic energy and, in the FP64 form, M; PART_C13: Cᵏ block columns 1 and 3 and,
// in the precision form, M; PART_C24: columns 2, 4).
// Every part evaluates the shared Gauss-point kinematics; the compiler keeps only what the part's
// outputs need, so each wavefront carries a fraction of the accumulators of the whole element.
enum { PART_STATIC = 0, PART_MASS = 1, PART_C13 = 2, PART_C24 = 3 };
// outputs a dynamic part can be asked for
enum { DYN_TK = 1, DYN_M = 2, DYN_C13 = 4, DYN_C24 = 8 };

// M blocks after l₀/2 Rₑ(·)Rₑᵀ and before the Tₛ⁻¹ post-multiplication, kept by the mass part
// for the damping term.
template <typename P> struct MassBlocks { M3<P> M11, M12, M13, M14, M22, M23, M24, M33, M34, M44; };

// Dynamic part (inertia, gyroscopic and damping forces, M and Cᵏ), compute_beams.jl:161-517.
template <typename P, typename S, int WHAT>
EBD void beam_dynamic(int64_t e, const NodeState& ns, const BeamConsts& bc, const BeamStage<S>& st,
                      const GaussPts& gp, const StaticOut& so, MassBlocks<P>& mb, S* mrec) {
    constexpr bool TK = (WHAT & DYN_TK) != 0, WM = (WHAT & DYN_M) != 0;
    constexpr bool CA = (WHAT & DYN_C13) != 0, CB = (WHAT & DYN_C24) != 0;
    const int64_t E = bc.E;
    const int32_t na = bc.n1[e], nb = bc.n2[e];
    const M3<P> I = eye<P>();
    const M3<P> Re = cvt<P>(so.Re);
    const M3<P> ReT = tr(Re);
    const P ln = P(so.ln), l0 = P(bc.l0[e]);
    const P Ar = P(bc.props[12 * E + e]), damping = P(bc.props[13 * E + e]);
    const M3<P> Jr = diag(P(bc.props[9 * E + e]), P(bc.props[10 * E + e]), P(bc.props[11 * E + e]));
    const V3<P> Tb1 = cvt<P>(so.Tb1), Tb2 = cvt<P>(so.Tb2);
    const M3<P> G1 = rows(P(0), P(0), P(so.eta / so.ln), P(0), P(0), P(1 / so.ln), P(0), P(-1 / so.ln), P(0));
    const M3<P> G2 = rows(P(so.eta12 / 2), P(-so.eta11 / 2), P(0), P(0), P(0), P(0), P(0), P(0), P(0));
    const M3<P> G3 = -G1;
    const M3<P> G4 = rows(P(so.eta22 / 2), P(-so.eta21 / 2), P(0), P(0), P(0), P(0), P(0), P(0), P(0));

    const V3<P> ud1 = cvt<P>(ldv(ns.ud, na)), ud2 = cvt<P>(ldv(ns.ud, nb));
    const V3<P> Ud1 = tmul(Re, ud1), Ud2 = tmul(Re, ud2);
    const V3<P> Wd1 = tmul(Re, cvt<P>(ldv(ns.wd, na))), Wd2 = tmul(Re, cvt<P>(ldv(ns.wd, nb)));
    const V3<P> Udd1 = tmul(Re, cvt<P>(ldv(ns.udd, na))), Udd2 = tmul(Re, cvt<P>(ldv(ns.udd, nb)));
    const V3<P> Wdd1 = tmul(Re, cvt<P>(ldv(ns.wdd, na))), Wdd2 = tmul(Re, cvt<P>(ldv(ns.wdd, nb)));
    const M3<P> SUd1 = skew(Ud1), SUd2 = skew(Ud2), SWd1 = skew(Wd1), SWd2 = skew(Wd2);
    const V3<P> We = ((G1 * Ud1 + G2 * Wd1) + G3 * Ud2) + G4 * Wd2;
    const M3<P> SWe = skew(We);
    const V3<P> v1 = cvt<P>(so.v1);
    const P rd = dot(-v1, ud1) + dot(v1, ud2);

    P kinetic = P(0);
    V3<P> Tk1 = zv<P>(), Tk2 = zv<P>(), Tk3 = zv<P>(), Tk4 = zv<P>();
    M3<P> M11 = zm<P>(), M12 = zm<P>(), M13 = zm<P>(), M14 = zm<P>(), M22 = zm<P>(), M23 = zm<P>(), M24 = zm<P>(),
          M33 = zm<P>(), M34 = zm<P>(), M44 = zm<P>();
    M3<P> C11 = zm<P>(), C12 = zm<P>(), C14 = zm<P>(), C21 = zm<P>(), C22 = zm<P>(), C24 = zm<P>(), C31 = zm<P>(),
          C32 = zm<P>(), C34 = zm<P>(), C41 = zm<P>(), C42 = zm<P>(), C44 = zm<P>();

#pragma unroll 1
    for (int ig = 0; ig < gp.n; ++ig) {
        const P w = P(gp.w[ig]);
        ShapeF sd = shape_fn(bc.l0[e], gp.z[ig]);
        const P N1 = P(sd.N1), N2 = P(sd.N2), N3 = P(sd.N3), N4 = P(sd.N4), N5 = P(sd.N5), N6 = P(sd.N6), N7 = P(sd.N7),
                N8 = P(sd.N8);
        V3<P> ut = {P(0), N3 * Tb1.z + N4 * Tb2.z, (-N3) * Tb1.y + (-N4) * Tb2.y};
        V3<P> Thb = {N1 * Tb1.x + N2 * Tb2.x, N5 * Tb1.y + N6 * Tb2.y, N5 * Tb1.z + N6 * Tb2.z};
        M3<P> Sut = skew(ut);
        M3<P> Rbar = I + skew(Thb);
        M3<P> Ir = (Rbar * Jr) * tr(Rbar);
        P N7l = N7 / ln, N7l2 = N7l / ln, N8l = N8 / ln, N8l2 = N8l / ln;
        M3<P> PP1 = rows(P(0), P(0), P(0), P(0), N7l, P(0), P(0), P(0), N7l);
        M3<P> PP2 = rows(P(0), P(0), P(0), P(0), P(0), N3, P(0), -N3, P(0));
        M3<P> PP3 = -PP1;
        M3<P> PP4 = rows(P(0), P(0), P(0), P(0), P(0), N4, P(0), -N4, P(0));
        M3<P> H11 = (N1 * I + PP1) - Sut * G1;
        M3<P> H12 = PP2 - Sut * G2;
        M3<P> H13 = (N2 * I + PP3) - Sut * G3;
        M3<P> H14 = PP4 - Sut * G4;
        M3<P> H21 = rows(P(0), P(0), P(0), P(0), P(0), -N8l, P(0), N8l, P(0));
        M3<P> H22 = diag(N1, N5, N5);
        M3<P> H23 = -H21;
        M3<P> H24 = diag(N2, N6, N6);
        V3<P> udt = ((PP1 * Ud1 + PP2 * Wd1) + PP3 * Ud2) + PP4 * Wd2;
        M3<P> Sudt = skew(udt);
        P N7rd = N7l2 * rd, N8rd = N8l2 * rd;
        M3<P> Hd11 = diag(P(0), -N7rd, -N7rd) - Sudt * G1;
        M3<P> Hd12 = -(Sudt * G2);
        M3<P> Hd14 = -(Sudt * G4);
        M3<P> Hd21 = rows(P(0), P(0), P(0), P(0), P(0), N8rd, P(0), -N8rd, P(0));
        V3<P> h1 = ((H11 * Ud1 + H12 * Wd1) + H13 * Ud2) + H14 * Wd2;
        V3<P> h2 = ((H21 * Ud1 + H22 * Wd1) + H23 * Ud2) + H24 * Wd2;
        M3<P> Sh1 = skew(h1), Sh2 = skew(h2);
        M3<P> C1_1 = (SWe * H11 + Hd11) - H11 * SWe;
        M3<P> C1_2 = (SWe * H12 + Hd12) - H12 * SWe;
        M3<P> C1_4 = (SWe * H14 + Hd14) - H14 * SWe;
        M3<P> C2_1 = (SWe * H21 + Hd21) - H21 * SWe;
        M3<P> C2_2 = SWe * H22 - H22 * SWe;
        M3<P> C2_4 = SWe * H24 - H24 * SWe;
        M3<P> H1F1 = ((H11 * SUd1 + H12 * SWd1) + H13 * SUd2) + H14 * SWd2;
        M3<P> H2F1 = ((H21 * SUd1 + H22 * SWd1) + H23 * SUd2) + H24 * SWd2;
        M3<P> A1 = rows(P(0), P(0), P(0), Ud1.y - Ud2.y, P(0), P(0), Ud1.z - Ud2.z, P(0), P(0));
        M3<P> C3_1 = (((-Sh1) * G1 + N7l2 * A1) + SWe * PP1) + H1F1 * G1;
        M3<P> C3_2 = ((-Sh1) * G2 + SWe * PP2) + H1F1 * G2;
        M3<P> C3_4 = ((-Sh1) * G4 + SWe * PP4) + H1F1 * G4;
        M3<P> A2 = rows(P(0), P(0), P(0), -Ud1.z + Ud2.z, P(0), P(0), Ud1.y - Ud2.y, P(0), P(0));
        M3<P> C4_1 = ((-Sh2) * G1 + N8l2 * A2) + H2F1 * G1;
        M3<P> C4_2 = (-Sh2) * G2 + H2F1 * G2;
        M3<P> C4_4 = (-Sh2) * G4 + H2F1 * G4;
        V3<P> ud0 = Re * h1;
        V3<P> ReTudd0 = (((H11 * Udd1 + H12 * Wdd1) + H13 * Udd2) + H14 * Wdd2) +
                        (((C1_1 * Ud1 + C1_2 * Wd1) + (-C1_1) * Ud2) + C1_4 * Wd2);
        V3<P> W0 = h2;
        V3<P> wd0 = Re * W0;
        V3<P> ReTwdd0 = (((H21 * Udd1 + H22 * Wdd1) + H23 * Udd2) + H24 * Wdd2) +
                        (((C2_1 * Ud1 + C2_2 * Wd1) + (-C2_1) * Ud2) + C2_4 * Wd2);
        M3<P> SW0Ir = skew(W0) * Ir;
        V3<P> X = Ir * ReTwdd0 + SW0Ir * W0;
        M3<P> AH11 = Ar * tr(H11), AH12 = Ar * tr(H12), AH14 = Ar * tr(H14);
        if constexpr (TK) {
        V3<P> Tk1G = w * (AH11 * ReTudd0 + tmul(H21, X));
        Tk1 += Tk1G;
        Tk2 += w * (AH12 * ReTudd0 + tmul(H22, X));
        Tk3 += (-Tk1G) + (w * Ar) * ReTudd0;
        Tk4 += w * (AH14 * ReTudd0 + tmul(H24, X));
        if (damping > P(0)) {
            V3<P> Irh2 = Ir * h2;
            V3<P> Td1G = (w * damping) * (AH11 * h1 + tmul(H21, Irh2));
            Tk1 += Td1G;
            Tk2 += (w * damping) * (AH12 * h1 + tmul(H22, Irh2));
            Tk3 += (-Td1G) + ((w * Ar) * damping) * h1;
            Tk4 += (w * damping) * (AH14 * h1 + tmul(H24, Irh2));
        }
        M3<P> Irg = (Re * Ir) * ReT;
        kinetic += (w / P(2)) * (Ar * dot(ud0, ud0) + dot(wd0, Irg * wd0));
        }
        if constexpr (WM) {
        M3<P> tH21Ir = tr(H21) * Ir, tH22Ir = tr(H22) * Ir, tH24Ir = tr(H24) * Ir;
        M3<P> M11G = w * (AH11 * H11 + tH21Ir * H21);
        M3<P> M12G = w * (AH11 * H12 + tH21Ir * H22);
        M3<P> M14G = w * (AH11 * H14 + tH21Ir * H24);
        M11 += M11G;
        M12 += M12G;
        M13 += (-M11G) + w * AH11;
        M14 += M14G;
        M22 += w * (AH12 * H12 + tH22Ir * H22);
        M23 += (-tr(M12G)) + w * AH12;
        M24 += w * (AH12 * H14 + tH22Ir * H24);
        M33 += M11G + (w * Ar) * ((I - H11) - tr(H11));
        M34 += (-M14G) + (w * Ar) * H14;
        M44 += w * (AH14 * H14 + tH24Ir * H24);
        }
        M3<P> SW0IrmS = SW0Ir - skew(Ir * W0);
        M3<P> AC1 = Ar * (C1_1 + C3_1), AC2 = Ar * (C1_2 + C3_2), AC4 = Ar * (C1_4 + C3_4);
        M3<P> Y1 = Ir * (C2_1 + C4_1) + SW0IrmS * H21;
        M3<P> Y2 = Ir * (C2_2 + C4_2) + SW0IrmS * H22;
        M3<P> Y4 = Ir * (C2_4 + C4_4) + SW0IrmS * H24;
        M3<P> tH11 = tr(H11), tH12 = tr(H12), tH14 = tr(H14), tH21 = tr(H21), tH22 = tr(H22), tH24 = tr(H24);
        if constexpr (CA) {
            M3<P> C11G = w * (tH11 * AC1 + tH21 * Y1);
            C11 += C11G;
            C21 += w * (tH12 * AC1 + tH22 * Y1);
            C31 += (-C11G) + w * AC1;
            C41 += w * (tH14 * AC1 + tH24 * Y1);
        }
        if constexpr (CB) {
            M3<P> C12G = w * (tH11 * AC2 + tH21 * Y2);
            M3<P> C14G = w * (tH11 * AC4 + tH21 * Y4);
            C12 += C12G; C14 += C14G;
            C22 += w * (tH12 * AC2 + tH22 * Y2);
            C24 += w * (tH12 * AC4 + tH22 * Y4);
            C32 += (-C12G) + w * AC2;
            C34 += (-C14G) + w * AC4;
            C42 += w * (tH14 * AC2 + tH24 * Y2);
            C44 += w * (tH14 * AC4 + tH24 * Y4);
        }
    }
    const P l02 = l0 / P(2);
    const M3<P> l02Re = l02 * Re;
    // exact Tₛ⁻¹(log ΔR)ᵀ post-multiplication of block rows 2 and 4 (compute_beams.jl:480-506)
    const M3<P> TT1t = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, na)))));
    const M3<P> TT2t = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, nb)))));
    if constexpr (TK) {
        st.energy[3 * e + 1] = double(l02 * kinetic);
        st_vec(st.Tk, E, e, 0, l02Re * Tk1);
        st_vec(st.Tk, E, e, 1, l02Re * Tk2);
        st_vec(st.Tk, E, e, 2, l02Re * Tk3);
        st_vec(st.Tk, E, e, 3, l02Re * Tk4);
    }
    if constexpr (WM) {
        M11 = (l02Re * M11) * ReT; M12 = (l02Re * M12) * ReT; M13 = (l02Re * M13) * ReT; M14 = (l02Re * M14) * ReT;
        M22 = (l02Re * M22) * ReT; M23 = (l02Re * M23) * ReT; M24 = (l02Re * M24) * ReT;
        M33 = (l02Re * M33) * ReT; M34 = (l02Re * M34) * ReT; M44 = (l02Re * M44) * ReT;
        mb = MassBlocks<P>{M11, M12, M13, M14, M22, M23, M24, M33, M34, M44};
        st_rec<S>(mrec, 0, 0, M11); st_rec<S>(mrec, 0, 1, M12);
        st_rec<S>(mrec, 0, 2, M13); st_rec<S>(mrec, 0, 3, M14);
        st_rec<S>(mrec, 1, 0, tr(M12) * TT1t);
        M3<P> M23t = M23 * TT1t, M24t = M24 * TT1t;
        st_rec<S>(mrec, 1, 1, M22 * TT1t); st_rec<S>(mrec, 1, 2, M23t); st_rec<S>(mrec, 1, 3, M24t);
        st_rec<S>(mrec, 2, 0, tr(M13)); st_rec<S>(mrec, 2, 1, tr(M23t));
        st_rec<S>(mrec, 2, 2, M33); st_rec<S>(mrec, 2, 3, M34);
        st_rec<S>(mrec, 3, 0, tr(M14) * TT2t); st_rec<S>(mrec, 3, 1, tr(M24t) * TT2t);
        st_rec<S>(mrec, 3, 2, tr(M34) * TT2t); st_rec<S>(mrec, 3, 3, M44 * TT2t);
    }
    if constexpr (CA) {
        C11 = (l02Re * C11) * ReT; C21 = (l02Re * C21) * ReT; C31 = (l02Re * C31) * ReT; C41 = (l02Re * C41) * ReT;
        st_block(st.C, E, e, 0, 0, C11); st_block(st.C, E, e, 0, 2, -C11);
        st_block(st.C, E, e, 1, 0, C21 * TT1t); st_block(st.C, E, e, 1, 2, (-C21) * TT1t);
        st_block(st.C, E, e, 2, 0, C31); st_block(st.C, E, e, 2, 2, -C31);
        st_block(st.C, E, e, 3, 0, C41 * TT2t); st_block(st.C, E, e, 3, 2, (-C41) * TT2t);
    }
    if constexpr (CB) {
        C12 = (l02Re * C12) * ReT; C14 = (l02Re * C14) * ReT; C22 = (l02Re * C22) * ReT; C24 = (l02Re * C24) * ReT;
        C32 = (l02Re * C32) * ReT; C34 = (l02Re * C34) * ReT; C42 = (l02Re * C42) * ReT; C44 = (l02Re * C44) * ReT;
        st_block(st.C, E, e, 0, 1, C12); st_block(st.C, E, e, 0, 3, C14);
        st_block(st.C, E, e, 1, 1, C22 * TT1t); st_block(st.C, E, e, 1, 3, C24 * TT1t);
        st_block(st.C, E, e, 2, 1, C32); st_block(st.C, E, e, 2, 3, C34);
        st_block(st.C, E, e, 3, 1, C42 * TT2t); st_block(st.C, E, e, 3, 3, C44 * TT2t);
    }
}

// Rayleigh-type damping term of Cᵏ (compute_beams.jl:462-479): Cᵏ += damping·M with the M blocks
// taken before the Tₛ⁻¹ post-multiplication, which then applies to the sum on block rows 2 and 4.
// Added by the mass part to the C blocks the other parts stored.
template <typename P, typename S>
EBD void beam_damping(int64_t e, const NodeState& ns, const BeamConsts& bc, const BeamStage<S>& st,
                      const MassBlocks<P>& m) {
    const int64_t E = bc.E;
    const P d = P(bc.props[13 * E + e]);
    const M3<P> TT1t = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, bc.n1[e])))));
    const M3<P> TT2t = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, bc.n2[e])))));
    auto add = [&](int bi, int bj, const M3<P>& B) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) st.C[stage_idx(e, 3 * bi + i, 3 * bj + j)] += S(B.a[i + 3 * j]);
    };
    add(0, 0, d * m.M11); add(0, 1, d * m.M12); add(0, 2, d * m.M13); add(0, 3, d * m.M14);
    add(1, 0, (d * tr(m.M12)) * TT1t); add(1, 1, (d * m.M22) * TT1t);
    add(1, 2, (d * m.M23) * TT1t); add(1, 3, (d * m.M24) * TT1t);
    add(2, 0, d * tr(m.M13)); add(2, 1, d * tr(m.M23)); add(2, 2, d * m.M33); add(2, 3, d * m.M34);
    add(3, 0, (d * tr(m.M14)) * TT2t); add(3, 1, (d * tr(m.M24)) * TT2t);
    add(3, 2, (d * tr(m.M34)) * TT2t); add(3, 3, (d * m.M44) * TT2t);
}

// One workgroup per 64 elements, four wavefronts: the static part and the three dynamic parts of
// the same 64 elements. Lane l of every wavefront works on element 64·blockIdx.x + l.
constexpr int BEAM_TILE = 64;

// S: storage type of the element matrices (double: the FP64 form; float: the precision form, whose
// Coriolis parts and the static part's tangent also compute in FP32). The static part's frame and
// Tⁱⁿᵗ compute in FP64, and the mass part too, because it yields the inertia forces Tᵏ: both enter
// the FP64 residual. A workgroup runs the NP parts P0 .. P0 + NP − 1 of its tile, one wavefront each
// (all four, or, in the precision form, the static and mass parts in one launch and the two
// Coriolis parts in the next).
template <typename S, int P0, int NP, typename Hook>
EBD void beam_kernel_body(int64_t tile, NodeState ns, BeamConsts bc, BeamStage<S> st, GaussPts gp,
                          const uint8_t* __restrict__ active, const int32_t* __restrict__ beam_inst, const Hook& hook) {
    constexpr bool HAS_S = P0 <= PART_STATIC && PART_STATIC < P0 + NP;
    constexpr bool HAS_M = P0 <= PART_MASS && PART_MASS < P0 + NP;
    constexpr bool HAS_13 = P0 <= PART_C13 && PART_C13 < P0 + NP;
    constexpr bool HAS_24 = P0 <= PART_C24 && PART_C24 < P0 + NP;
    static_assert(!Hook::SPLIT || (HAS_S && HAS_M), "the contact test is shared by the static and mass parts");
    __shared__ S k_lds[BEAM_TILE * STAGE_REC];
    __shared__ S m_lds[BEAM_TILE * STAGE_REC];
    __shared__ uint8_t hook_bits[BEAM_TILE];
    const int part = P0 + __builtin_amdgcn_readfirstlane(threadIdx.x / BEAM_TILE);
    const int lane = threadIdx.x % BEAM_TILE;
    const int64_t e = tile * BEAM_TILE + lane;
    // elements of instances frozen in the Newton iteration are skipped
    const bool valid = e < bc.E && active[beam_inst[e]];
    const uint64_t vmask = __ballot(valid);   // the same 64 elements in every wavefront
    if (vmask == 0) return;                   // the whole workgroup: a tile of frozen instances
    // FP64 form: the mass part computes M with Tᵏ; precision form: Tᵏ alone in FP64 (it enters the
    // residual), M in FP32 with the Coriolis columns 1 and 3
    constexpr bool FP64 = sizeof(S) == sizeof(double);
    constexpr int MASS_WHAT = FP64 ? (DYN_TK | DYN_M) : DYN_TK;
    constexpr int C13_WHAT = FP64 ? DYN_C13 : (DYN_C13 | DYN_M);
    const int m_part = FP64 ? PART_MASS : PART_C13;
    constexpr bool HAS_MP = FP64 ? HAS_M : HAS_13;   // the workgroup holds the part computing M
    // the work beside the element (the contact test of its Gauss points) is shared by the two parts
    // with the least work of their own: the static part takes the middle point, the mass part the others
    const unsigned all_pts = (1u << gp.n) - 1;
    const unsigned static_pts = gp.n >= 2 ? 2u : 0u;
    MassBlocks<double> mbd;
    MassBlocks<S> mbs;
    unsigned own = 0;
    if (valid) {
        if (HAS_S && part == PART_STATIC) {
            beam_static<S>(e, ns, bc, st, k_lds + lane * STAGE_REC,
                           [&](const V3<double>& x1, const V3<double>& x2, const M3<double>& Re, const V3<double>& Tb1,
                               const V3<double>& Tb2) {
                               if constexpr (Hook::SPLIT)
                                   hook_bits[lane] = (uint8_t)hook.test(e, bc, gp, static_pts, x1, x2, Re, Tb1, Tb2);
                           });
        } else if constexpr (HAS_M || HAS_13 || HAS_24) {
            const StaticOut so = beam_frame(e, ns, bc);
            if (HAS_M && part == PART_MASS) {
                if constexpr (Hook::SPLIT) {
                    const int32_t a = bc.n1[e], b = bc.n2[e];
                    own = hook.test(e, bc, gp, all_pts & ~static_pts, ldv(ns.X0, a) + ldv(ns.u, a),
                                    ldv(ns.X0, b) + ldv(ns.u, b), so.Re, so.Tb1, so.Tb2);
                }
                beam_dynamic<double, S, MASS_WHAT>(e, ns, bc, st, gp, so, mbd, m_lds + lane * STAGE_REC);
            } else if (HAS_13 && part == PART_C13) {
                beam_dynamic<S, S, C13_WHAT>(e, ns, bc, st, gp, so, mbs, m_lds + lane * STAGE_REC);
            } else if constexpr (HAS_24) {
                beam_dynamic<S, S, DYN_C24>(e, ns, bc, st, gp, so, mbs, nullptr);
            }
        }
    }
    // each wavefront copies out the records it staged
    if (HAS_S && part == PART_STATIC) tile_records_out(st.K + tile * BEAM_TILE * 144, k_lds, vmask, lane);
    if (HAS_MP && part == m_part) tile_records_out(st.M + tile * BEAM_TILE * 144, m_lds, vmask, lane);
    // every wavefront sees the same 64 elements, so this test is uniform over the workgroup
    const bool damped = __ballot(valid && bc.props[13 * bc.E + e] > 0.0) != 0;
    if (damped) {
        __syncthreads();
        // (Cᵏ is complete here only when the workgroup holds both Coriolis parts: the four-part
        // kernel, or the second launch of the precision form, where M is computed too)
        if (HAS_MP && HAS_13 && HAS_24 && part == m_part && valid && bc.props[13 * bc.E + e] > 0.0) {
            if constexpr (FP64) beam_damping<double, S>(e, ns, bc, st, mbd);
            else beam_damping<S, S>(e, ns, bc, st, mbs);
        }
    }
    if constexpr (Hook::SPLIT) {
        __syncthreads();   // the static part's test is in LDS
        if (part == PART_MASS && valid) hook.finish(e, own | hook_bits[lane]);
    }
}

template <typename S>
__global__ void __launch_bounds__(4 * BEAM_TILE) beam_kernel(NodeState ns, BeamConsts bc, BeamStage<S> st, GaussPts gp,
                                                              const uint8_t* __restrict__ active,
                                                              const int32_t* __restrict__ beam_inst) {
    beam_kernel_body<S, 0, 4>(blockIdx.x, ns, bc, st, gp, active, beam_inst, NoStaticHook{});
}
// Two of the four parts per workgroup (P0 = PART_STATIC: static and mass; P0 = PART_C13: the Coriolis
// parts): the precision form's element kernel as two launches. Each launch's code is about half
// the four-part kernel's. The static and mass parts need about 320 VGPRs (one wavefront per SIMD:
// a CU holds two of their workgroups); the Coriolis parts are held to 256 (10 spilled), two
// wavefronts per SIMD, four workgroups per CU.
template <typename S, int P0>
__global__ void __launch_bounds__(2 * BEAM_TILE, P0 == PART_C13 ? 2 : 1) beam_kernel_half(NodeState ns, BeamConsts bc, BeamStage<S> st,
                                                                   GaussPts gp, const uint8_t* __restrict__ active,
                                                                   const int32_t* __restrict__ beam_inst) {
    beam_kernel_body<S, P0, 2>(blockIdx.x, ns, bc, st, gp, active, beam_inst, NoStaticHook{});
}

}  // namespace ebd
