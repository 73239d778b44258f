// eb_elements.h — TEST INFRASTRUCTURE (oracle). Restates the reference's per-element kernels:
//   compute_beams          src/beams/compute_beams.jl:1-538
//   compute_contact_beams  src/interactions/compute&assemble_interactions_beams.jl:1-445
//   gap functions          src/interactions/utils_compute_interactions_beams.jl:1-75
//   regularisation         src/interactions/interactions_beams_regularization.jl:1-45
// statement by statement, in FP64, with the evaluation order documented in eb_math.h.
#pragma once
#include "eb_math.h"
#include <cstdint>
#include <vector>

namespace ebo {

struct BeamProps {  // BeamMaterialProperties (src/beams/def_beam.jl:19-26)
    real radius, E, Ku;
    V3 KT, KTT, Jrho;
    real Arho, damping;
};
inline BeamProps props_from(const real* p) {
    BeamProps b;
    b.radius = p[0]; b.E = p[1]; b.Ku = p[2];
    b.KT = v3(p[3], p[4], p[5]); b.KTT = v3(p[6], p[7], p[8]); b.Jrho = v3(p[9], p[10], p[11]);
    b.Arho = p[12]; b.damping = p[13];
    return b;
}

struct Gauss { int n; real w[8], z[8]; };

struct ElemIn {
    V3 X1, X2, u1, u2, ud1, ud2, udd1, udd2, wd1, wd2, wdd1, wdd2;
    M3 R1, R2, dR1, dR2;
    real l0;
    M3 Re0;
};

// 12x12 column-major, local dof order [u₁ θ₁ u₂ θ₂]
struct Mat12 {
    real a[144];
    void set_block(int bi, int bj, const M3& B) {
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) a[(3 * bj + j) * 12 + 3 * bi + i] = B(i, j);
    }
};

struct BeamOut {
    real strain, kinetic;
    real Tint[12], Tk[12];
    Mat12 K, M, C;
};

inline void set_vec(real* t, int b, const V3& v) { t[3 * b] = v[0]; t[3 * b + 1] = v[1]; t[3 * b + 2] = v[2]; }

// shape functions of compute_beams.jl:238-248 (identical in the contact kernel :58-68)
struct Shape { real N1, N2, N3, N4, N5, N6, N7, N8; };
inline Shape shape(real l0, real z) {
    Shape s;
    real xi = l0 * (z + 1) / 2;
    s.N1 = 1 - xi / l0;
    s.N2 = 1 - s.N1;
    real a = 1 - xi / l0;
    s.N3 = xi * (a * a);
    s.N4 = -(1 - xi / l0) * ((xi * xi) / l0);
    s.N5 = (1 - 3 * xi / l0) * (1 - xi / l0);
    s.N6 = (3 * xi / l0 - 2) * (xi / l0);
    s.N7 = s.N3 + s.N4;
    s.N8 = s.N5 + s.N6 - 1;
    return s;
}

// compute_beams(u₁,…,constants, exact=true, isdynamic=true), src/beams/compute_beams.jl:1-538
inline void compute_beams(const ElemIn& in, const BeamProps& bp, const Gauss& g, BeamOut& out) {
    const M3 I = eye();
    V3 x1 = in.X1 + in.u1, x2 = in.X2 + in.u2;
    real ln = norm(x2 - x1);
    real ub = ln - in.l0;
    LocalFrame f = local_Re_and_aux(x1, x2, in.R1, in.R2, col(in.Re0, 1), ln);
    const M3& Re = f.Re;
    M3 ReT = tr(Re);
    const M3 &G1 = f.G1, &G2 = f.G2, &G3 = f.G3, &G4 = f.G4;
    M3 Rb1 = (ReT * in.R1) * in.Re0;
    M3 Rb2 = (ReT * in.R2) * in.Re0;
    V3 Tb1 = toangle(Rb1), Tb2 = toangle(Rb2);
    M3 Ti1 = Ts_inv(Tb1), Ti2 = Ts_inv(Tb2);

    M3 P11 = -G1, P21 = P11, P12 = I - G2, P22 = -G2, P14 = -G4, P24 = I - G4;
    V3 Bp1 = f.r1;  // row vector r¹
    M3 Bp11 = P11 * ReT, Bp21 = Bp11, Bp12 = P12 * ReT, Bp22 = P22 * ReT, Bp14 = P14 * ReT, Bp24 = P24 * ReT;
    V3 B1 = Bp1;
    M3 B11 = Ti1 * Bp11, B12 = Ti1 * Bp12, B14 = Ti1 * Bp14;
    M3 B21 = Ti2 * Bp21, B22 = Ti2 * Bp22, B24 = Ti2 * Bp24;

    real Ku = bp.Ku;
    M3 KT = diag(bp.KT[0], bp.KT[1], bp.KT[2]), KTT = diag(bp.KTT[0], bp.KTT[1], bp.KTT[2]);
    real Tu = Ku * ub;
    V3 TT1 = KT * Tb1 + KTT * Tb2;
    V3 TT2 = KTT * Tb1 + KT * Tb2;
    out.strain = (ub * Tu + dot(Tb1, TT1) + dot(Tb2, TT2)) / 2;

    V3 T1 = (B1 * Tu + tr(B11) * TT1) + tr(B21) * TT2;
    V3 T2 = tr(B12) * TT1 + tr(B22) * TT2;
    V3 T3 = -T1;
    V3 T4 = tr(B14) * TT1 + tr(B24) * TT2;
    set_vec(out.Tint, 0, T1); set_vec(out.Tint, 1, T2); set_vec(out.Tint, 2, T3); set_vec(out.Tint, 3, T4);

    real Nb = Tu;
    V3 Mp1 = tr(Ti1) * TT1, Mp2 = tr(Ti2) * TT2;
    V3 Qs1 = tr(P11) * Mp1 + tr(P21) * Mp2;
    V3 Qs2 = tr(P12) * Mp1 + tr(P22) * Mp2;
    V3 Qs4 = tr(P14) * Mp1 + tr(P24) * Mp2;
    M3 Q1 = skew(Qs1), Q2 = skew(Qs2), Q4 = skew(Qs4);
    V3 a = v3(0, f.eta * (Mp1[0] + Mp2[0]) / ln + (Mp1[1] + Mp2[1]) / ln, (Mp1[2] + Mp2[2]) / ln);
    M3 DN11 = f.D3 * Nb;
    M3 QG11 = Q1 * G1, QG12 = Q1 * G2, QG14 = Q1 * G4, QG22 = Q2 * G2, QG24 = Q2 * G4, QG44 = Q4 * G4;
    V3 EGa1 = Re * (tr(G1) * a);
    M3 EGar11 = outer(EGa1, f.r1);
    M3 Km11 = (DN11 - (Re * QG11) * ReT) + EGar11;
    M3 Km12 = -((Re * QG12) * ReT);
    M3 Km14 = -((Re * QG14) * ReT);
    M3 Km22 = -((Re * QG22) * ReT);
    M3 Km24 = -((Re * QG24) * ReT);
    M3 Km44 = -((Re * QG44) * ReT);

    real eta1, mu1, eta2, mu2;
    eta_mu(Tb1, eta1, mu1);
    eta_mu(Tb2, eta2, mu2);
    M3 Kh1 = Kh_bar(Tb1, TT1, Ti1, eta1, mu1);
    M3 Kh2 = Kh_bar(Tb2, TT2, Ti2, eta2, mu2);
    M3 Kt11 = ((tr(Bp11) * Kh1) * Bp11 + (tr(Bp21) * Kh2) * Bp21) + Km11;
    M3 Kt12 = ((tr(Bp11) * Kh1) * Bp12 + (tr(Bp21) * Kh2) * Bp22) + Km12;
    M3 Kt14 = ((tr(Bp11) * Kh1) * Bp14 + (tr(Bp21) * Kh2) * Bp24) + Km14;
    M3 Kt22 = ((tr(Bp12) * Kh1) * Bp12 + (tr(Bp22) * Kh2) * Bp22) + Km22;
    M3 Kt24 = ((tr(Bp12) * Kh1) * Bp14 + (tr(Bp22) * Kh2) * Bp24) + Km24;
    M3 Kt44 = ((tr(Bp14) * Kh1) * Bp14 + (tr(Bp24) * Kh2) * Bp24) + Km44;

    M3 BKB11 = (outer(B1 * Ku, B1) + tr(B11) * (KT * B11 + KTT * B21)) + tr(B21) * (KTT * B11 + KT * B21);
    M3 BKB12 = tr(B11) * (KT * B12 + KTT * B22) + tr(B21) * (KTT * B12 + KT * B22);
    M3 BKB14 = tr(B11) * (KT * B14 + KTT * B24) + tr(B21) * (KTT * B14 + KT * B24);
    M3 BKB22 = tr(B12) * (KT * B12 + KTT * B22) + tr(B22) * (KTT * B12 + KT * B22);
    M3 BKB24 = tr(B12) * (KT * B14 + KTT * B24) + tr(B22) * (KTT * B14 + KT * B24);
    M3 BKB44 = tr(B14) * (KT * B14 + KTT * B24) + tr(B24) * (KTT * B14 + KT * B24);

    M3 K11 = BKB11 + Kt11, K12 = BKB12 + Kt12, K13 = -K11, K14 = BKB14 + Kt14;
    M3 K22 = BKB22 + Kt22, K23 = -tr(K12), K24 = BKB24 + Kt24;
    M3 K33 = K11, K34 = -K14, K44 = BKB44 + Kt44;
    // hcat(vcat(K¹¹, K¹²', K¹³', K¹⁴'), vcat(K¹², K²², K²³', K²⁴'), vcat(K¹³, K²³, K³³, K³⁴'), vcat(K¹⁴, K²⁴, K³⁴, K⁴⁴))
    out.K.set_block(0, 0, K11); out.K.set_block(1, 0, tr(K12)); out.K.set_block(2, 0, tr(K13)); out.K.set_block(3, 0, tr(K14));
    out.K.set_block(0, 1, K12); out.K.set_block(1, 1, K22); out.K.set_block(2, 1, tr(K23)); out.K.set_block(3, 1, tr(K24));
    out.K.set_block(0, 2, K13); out.K.set_block(1, 2, K23); out.K.set_block(2, 2, K33); out.K.set_block(3, 2, tr(K34));
    out.K.set_block(0, 3, K14); out.K.set_block(1, 3, K24); out.K.set_block(2, 3, K34); out.K.set_block(3, 3, K44);

    // ---- dynamic part, compute_beams.jl:161-517
    real kinetic = 0.0;
    V3 Tk1 = zv(), Tk2 = zv(), Tk3 = zv(), Tk4 = zv();
    M3 M11 = zm(), M12 = zm(), M13 = zm(), M14 = zm(), M22 = zm(), M23 = zm(), M24 = zm(), M33 = zm(), M34 = zm(), M44 = zm();
    M3 C11 = zm(), C12 = zm(), C14 = zm(), C21 = zm(), C22 = zm(), C24 = zm(), C31 = zm(), C32 = zm(), C34 = zm(),
       C41 = zm(), C42 = zm(), C44 = zm();
    const real Ar = bp.Arho, damping = bp.damping;
    M3 Jr = diag(bp.Jrho[0], bp.Jrho[1], bp.Jrho[2]);

    V3 Ud1 = ReT * in.ud1, Ud2 = ReT * in.ud2, Wd1 = ReT * in.wd1, Wd2 = ReT * in.wd2;
    V3 Udd1 = ReT * in.udd1, Udd2 = ReT * in.udd2, Wdd1 = ReT * in.wdd1, Wdd2 = ReT * in.wdd2;
    M3 SUd1 = skew(Ud1), SUd2 = skew(Ud2), SWd1 = skew(Wd1), SWd2 = skew(Wd2);
    V3 We = ((G1 * Ud1 + G2 * Wd1) + G3 * Ud2) + G4 * Wd2;
    M3 SWe = skew(We);
    real rd = dot(f.r1, in.ud1) + dot(f.r3, in.ud2);

    for (int ig = 0; ig < g.n; ++ig) {
        real z = g.z[ig], w = g.w[ig];
        Shape s = shape(in.l0, z);
        V3 ut = v3(0, s.N3 * Tb1[2] + s.N4 * Tb2[2], (-s.N3) * Tb1[1] + (-s.N4) * Tb2[1]);
        V3 Thb = v3(s.N1 * Tb1[0] + s.N2 * Tb2[0], s.N5 * Tb1[1] + s.N6 * Tb2[1], s.N5 * Tb1[2] + s.N6 * Tb2[2]);
        M3 Sut = skew(ut), STh = skew(Thb);
        M3 Rbar = I + STh;
        M3 Ir = (Rbar * Jr) * tr(Rbar);
        real N7l = s.N7 / ln, N7l2 = N7l / ln, N8l = s.N8 / ln, N8l2 = N8l / ln;
        M3 PP1 = rows(0, 0, 0, 0, N7l, 0, 0, 0, N7l);
        M3 PP2 = rows(0, 0, 0, 0, 0, s.N3, 0, -s.N3, 0);
        M3 PP3 = -PP1;
        M3 PP4 = rows(0, 0, 0, 0, 0, s.N4, 0, -s.N4, 0);
        M3 H11 = (s.N1 * I + PP1) - Sut * G1;
        M3 H12 = PP2 - Sut * G2;
        M3 H13 = (s.N2 * I + PP3) - Sut * G3;
        M3 H14 = PP4 - Sut * G4;
        M3 H21 = rows(0, 0, 0, 0, 0, -N8l, 0, N8l, 0);
        M3 H22 = diag(s.N1, s.N5, s.N5);
        M3 H23 = -H21;
        M3 H24 = diag(s.N2, s.N6, s.N6);
        V3 udt = ((PP1 * Ud1 + PP2 * Wd1) + PP3 * Ud2) + PP4 * Wd2;
        M3 Sudt = skew(udt);
        real N7rd = N7l2 * rd;
        M3 Hd11 = diag(0, -N7rd, -N7rd) - Sudt * G1;
        M3 Hd12 = -(Sudt * G2);
        M3 Hd14 = -(Sudt * G4);
        real N8rd = N8l2 * rd;
        M3 Hd21 = rows(0, 0, 0, 0, 0, N8rd, 0, -N8rd, 0);
        V3 h1 = ((H11 * Ud1 + H12 * Wd1) + H13 * Ud2) + H14 * Wd2;
        V3 h2 = ((H21 * Ud1 + H22 * Wd1) + H23 * Ud2) + H24 * Wd2;
        M3 Sh1 = skew(h1), Sh2 = skew(h2);
        M3 C1_1 = (SWe * H11 + Hd11) - H11 * SWe;
        M3 C1_2 = (SWe * H12 + Hd12) - H12 * SWe;
        M3 C1_3 = -C1_1;
        M3 C1_4 = (SWe * H14 + Hd14) - H14 * SWe;
        M3 C2_1 = (SWe * H21 + Hd21) - H21 * SWe;
        M3 C2_2 = SWe * H22 - H22 * SWe;
        M3 C2_3 = -C2_1;
        M3 C2_4 = SWe * H24 - H24 * SWe;
        M3 H1F1 = ((H11 * SUd1 + H12 * SWd1) + H13 * SUd2) + H14 * SWd2;
        M3 H2F1 = ((H21 * SUd1 + H22 * SWd1) + H23 * SUd2) + H24 * SWd2;
        M3 A1 = rows(0, 0, 0, Ud1[1] - Ud2[1], 0, 0, Ud1[2] - Ud2[2], 0, 0);
        M3 C3_1 = (((-Sh1) * G1 + N7l2 * A1) + SWe * PP1) + H1F1 * G1;
        M3 C3_2 = ((-Sh1) * G2 + SWe * PP2) + H1F1 * G2;
        M3 C3_4 = ((-Sh1) * G4 + SWe * PP4) + H1F1 * G4;
        M3 A2 = rows(0, 0, 0, -Ud1[2] + Ud2[2], 0, 0, Ud1[1] - Ud2[1], 0, 0);
        M3 C4_1 = ((-Sh2) * G1 + N8l2 * A2) + H2F1 * G1;
        M3 C4_2 = (-Sh2) * G2 + H2F1 * G2;
        M3 C4_4 = (-Sh2) * G4 + H2F1 * G4;
        V3 ud0 = Re * h1;
        V3 H1Edd = ((H11 * Udd1 + H12 * Wdd1) + H13 * Udd2) + H14 * Wdd2;
        V3 C1Ed = ((C1_1 * Ud1 + C1_2 * Wd1) + C1_3 * Ud2) + C1_4 * Wd2;
        V3 ReTudd0 = H1Edd + C1Ed;
        V3 W0 = h2;
        V3 wd0 = Re * W0;
        V3 H2Edd = ((H21 * Udd1 + H22 * Wdd1) + H23 * Udd2) + H24 * Wdd2;
        V3 C2Ed = ((C2_1 * Ud1 + C2_2 * Wd1) + C2_3 * Ud2) + C2_4 * Wd2;
        V3 ReTwdd0 = H2Edd + C2Ed;
        M3 SW0 = skew(W0);
        V3 IrRwdd = Ir * ReTwdd0;
        M3 SW0Ir = SW0 * Ir;
        V3 SW0IrW0 = SW0Ir * W0;
        V3 X = IrRwdd + SW0IrW0;
        M3 AH11 = Ar * tr(H11), AH12 = Ar * tr(H12), AH13 = Ar * tr(H13), AH14 = Ar * tr(H14);
        (void)AH13;
        V3 Tk1G = w * (AH11 * ReTudd0 + tr(H21) * X);
        Tk1 += Tk1G;
        Tk2 += w * (AH12 * ReTudd0 + tr(H22) * X);
        Tk3 += (-Tk1G) + (w * Ar) * ReTudd0;
        Tk4 += w * (AH14 * ReTudd0 + tr(H24) * X);
        if (damping > 0) {
            V3 Td1G = (w * damping) * (AH11 * h1 + tr(H21) * (Ir * h2));
            Tk1 += Td1G;
            Tk2 += (w * damping) * (AH12 * h1 + tr(H22) * (Ir * h2));
            Tk3 += (-Td1G) + ((w * Ar) * damping) * h1;
            Tk4 += (w * damping) * (AH14 * h1 + tr(H24) * (Ir * h2));
        }
        M3 M11G = w * (AH11 * H11 + (tr(H21) * Ir) * H21);
        M3 M12G = w * (AH11 * H12 + (tr(H21) * Ir) * H22);
        M3 M14G = w * (AH11 * H14 + (tr(H21) * Ir) * H24);
        M3 M22G = w * (AH12 * H12 + (tr(H22) * Ir) * H22);
        M3 M24G = w * (AH12 * H14 + (tr(H22) * Ir) * H24);
        M3 M44G = w * (AH14 * H14 + (tr(H24) * Ir) * H24);
        M11 += M11G;
        M12 += M12G;
        M13 += (-M11G) + w * AH11;
        M14 += M14G;
        M22 += M22G;
        M23 += (-tr(M12G)) + w * AH12;
        M24 += M24G;
        M33 += M11G + (w * Ar) * ((I - H11) - tr(H11));
        M34 += (-M14G) + (w * Ar) * H14;
        M44 += M44G;
        M3 SW0IrmS = SW0Ir - skew(Ir * W0);
        M3 AC1 = Ar * (C1_1 + C3_1), AC2 = Ar * (C1_2 + C3_2), AC4 = Ar * (C1_4 + C3_4);
        M3 Y1 = Ir * (C2_1 + C4_1) + SW0IrmS * H21;
        M3 Y2 = Ir * (C2_2 + C4_2) + SW0IrmS * H22;
        M3 Y4 = Ir * (C2_4 + C4_4) + SW0IrmS * H24;
        M3 C11G = w * (tr(H11) * AC1 + tr(H21) * Y1);
        M3 C12G = w * (tr(H11) * AC2 + tr(H21) * Y2);
        M3 C14G = w * (tr(H11) * AC4 + tr(H21) * Y4);
        M3 C21G = w * (tr(H12) * AC1 + tr(H22) * Y1);
        M3 C22G = w * (tr(H12) * AC2 + tr(H22) * Y2);
        M3 C24G = w * (tr(H12) * AC4 + tr(H22) * Y4);
        M3 C41G = w * (tr(H14) * AC1 + tr(H24) * Y1);
        M3 C42G = w * (tr(H14) * AC2 + tr(H24) * Y2);
        M3 C44G = w * (tr(H14) * AC4 + tr(H24) * Y4);
        C11 += C11G; C12 += C12G; C14 += C14G;
        C21 += C21G; C22 += C22G; C24 += C24G;
        C31 += (-C11G) + w * AC1;
        C32 += (-C12G) + w * AC2;
        C34 += (-C14G) + w * AC4;
        C41 += C41G; C42 += C42G; C44 += C44G;
        M3 Irg = (Re * Ir) * ReT;
        kinetic += (w / 2) * (Ar * dot(ud0, ud0) + dot(wd0, Irg * wd0));
    }
    real l02 = in.l0 / 2;
    M3 l02Re = l02 * Re;
    Tk1 = l02Re * Tk1; Tk2 = l02Re * Tk2; Tk3 = l02Re * Tk3; Tk4 = l02Re * Tk4;
    M11 = (l02Re * M11) * ReT; M12 = (l02Re * M12) * ReT; M13 = (l02Re * M13) * ReT; M14 = (l02Re * M14) * ReT;
    M22 = (l02Re * M22) * ReT; M23 = (l02Re * M23) * ReT; M24 = (l02Re * M24) * ReT;
    M33 = (l02Re * M33) * ReT; M34 = (l02Re * M34) * ReT; M44 = (l02Re * M44) * ReT;
    C11 = (l02Re * C11) * ReT; C12 = (l02Re * C12) * ReT; M3 C13 = -C11; C14 = (l02Re * C14) * ReT;
    C21 = (l02Re * C21) * ReT; C22 = (l02Re * C22) * ReT; M3 C23 = -C21; C24 = (l02Re * C24) * ReT;
    C31 = (l02Re * C31) * ReT; C32 = (l02Re * C32) * ReT; M3 C33 = -C31; C34 = (l02Re * C34) * ReT;
    C41 = (l02Re * C41) * ReT; C42 = (l02Re * C42) * ReT; M3 C43 = -C41; C44 = (l02Re * C44) * ReT;
    out.kinetic = l02 * kinetic;
    if (damping > 0) {
        C11 += damping * M11; C12 += damping * M12; C13 += damping * M13; C14 += damping * M14;
        C21 += damping * tr(M12); C22 += damping * M22; C23 += damping * M23; C24 += damping * M24;
        C31 += damping * tr(M13); C32 += damping * tr(M23); C33 += damping * M33; C34 += damping * M34;
        C41 += damping * tr(M14); C42 += damping * tr(M24); C43 += damping * tr(M34); C44 += damping * M44;
    }
    // exact: right-multiplication of block rows 2 and 4 by Tₛ⁻¹(log ΔR)ᵀ, compute_beams.jl:480-506.
    // Sequential in place: M³² uses the already transformed M²³, M⁴² the transformed M²⁴.
    V3 Th1 = toangle(in.dR1), Th2 = toangle(in.dR2);
    M3 TT1t = tr(Ts_inv(Th1)), TT2t = tr(Ts_inv(Th2));
    M3 M21 = tr(M12) * TT1t;
    M22 = M22 * TT1t;
    M23 = M23 * TT1t;
    M24 = M24 * TT1t;
    M3 M31 = tr(M13);
    M3 M32 = tr(M23);
    M3 M41 = tr(M14) * TT2t;
    M3 M42 = tr(M24) * TT2t;
    M3 M43 = tr(M34) * TT2t;
    M44 = M44 * TT2t;
    C21 = C21 * TT1t; C22 = C22 * TT1t; C23 = C23 * TT1t; C24 = C24 * TT1t;
    C41 = C41 * TT2t; C42 = C42 * TT2t; C43 = C43 * TT2t; C44 = C44 * TT2t;

    set_vec(out.Tk, 0, Tk1); set_vec(out.Tk, 1, Tk2); set_vec(out.Tk, 2, Tk3); set_vec(out.Tk, 3, Tk4);
    const M3* Mb[16] = {&M11, &M12, &M13, &M14, &M21, &M22, &M23, &M24, &M31, &M32, &M33, &M34, &M41, &M42, &M43, &M44};
    const M3* Cb[16] = {&C11, &C12, &C13, &C14, &C21, &C22, &C23, &C24, &C31, &C32, &C33, &C34, &C41, &C42, &C43, &C44};
    for (int bi = 0; bi < 4; ++bi)
        for (int bj = 0; bj < 4; ++bj) {
            out.M.set_block(bi, bj, *Mb[4 * bi + bj]);
            out.C.set_block(bi, bj, *Cb[4 * bi + bj]);
        }
}

// ---- contact -------------------------------------------------------------------------------

struct ContactProps { real kn, mu, eps_t, eta_n; };

// quadratic B-spline SDF, Interpolations.jl BSpline(Quadratic(Reflect(OnCell()))) + scale
// (src/interactions/def_interactions.jl:67-82). Coefficients padded by one cell on each side.
struct SDF {
    int64_t n[3];
    real d[3], x0[3];
    std::vector<real> coef;  // (n0+2)*(n1+2)*(n2+2), x fastest
    real dom[6];
    real c(int64_t i, int64_t j, int64_t k) const {
        return coef[(size_t)((k * (n[1] + 2) + j) * (n[0] + 2) + i)];
    }
};

// in-place solve of the (n+2)-point system  c₀ = c₁,  c_{i-1}/8 + 3c_i/4 + c_{i+1}/8 = y_i,
// c_{n+1} = c_n  along one line (Thomas algorithm on the reduced n×n system)
inline void prefilter_line(const real* y, real* c, int64_t n, std::vector<real>& cp, std::vector<real>& dp) {
    if (n == 1) { c[1] = y[0]; c[0] = c[1]; c[2] = c[1]; return; }
    cp.resize(n); dp.resize(n);
    // row 0: (3/4 + 1/8) c1 + 1/8 c2 = y0 ; row n-1: 1/8 c_{n-1} + (3/4+1/8) c_n = y_{n-1}
    for (int64_t i = 0; i < n; ++i) {
        real a = (i == 0) ? 0.0 : 0.125;
        real b = (i == 0 || i == n - 1) ? 0.875 : 0.75;
        real cc = (i == n - 1) ? 0.0 : 0.125;
        real den = b - (i > 0 ? a * cp[i - 1] : 0.0);
        cp[i] = cc / den;
        dp[i] = (y[i] - (i > 0 ? a * dp[i - 1] : 0.0)) / den;
    }
    c[n] = dp[n - 1];
    for (int64_t i = n - 2; i >= 0; --i) c[i + 1] = dp[i] - cp[i] * c[i + 2];
    c[0] = c[1];
    c[n + 1] = c[n];
}

inline void sdf_build(SDF& s, const int64_t n[3], const real d[3], const real x0[3], const real* samples) {
    for (int a = 0; a < 3; ++a) { s.n[a] = n[a]; s.d[a] = d[a]; s.x0[a] = x0[a]; }
    s.dom[0] = x0[0]; s.dom[1] = x0[0] + (n[0] - 1) * d[0];
    s.dom[2] = x0[1]; s.dom[3] = x0[1] + (n[1] - 1) * d[1];
    s.dom[4] = x0[2]; s.dom[5] = x0[2] + (n[2] - 1) * d[2];
    int64_t P0 = n[0] + 2, P1 = n[1] + 2, P2 = n[2] + 2;
    s.coef.assign((size_t)(P0 * P1 * P2), 0.0);
    auto at = [&](int64_t i, int64_t j, int64_t k) -> real& { return s.coef[(size_t)((k * P1 + j) * P0 + i)]; };
    std::vector<real> y, c, cp, dp;
    // x lines (interior j,k)
    y.resize(n[0]); c.resize(n[0] + 2);
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j) {
            for (int64_t i = 0; i < n[0]; ++i) y[i] = samples[(k * n[1] + j) * n[0] + i];
            prefilter_line(y.data(), c.data(), n[0], cp, dp);
            for (int64_t i = 0; i < P0; ++i) at(i, j + 1, k + 1) = c[i];
        }
    // y lines for every padded i
    y.resize(n[1]); c.resize(n[1] + 2);
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t i = 0; i < P0; ++i) {
            for (int64_t j = 0; j < n[1]; ++j) y[j] = at(i, j + 1, k + 1);
            prefilter_line(y.data(), c.data(), n[1], cp, dp);
            for (int64_t j = 0; j < P1; ++j) at(i, j, k + 1) = c[j];
        }
    // z lines for every padded i, j
    y.resize(n[2]); c.resize(n[2] + 2);
    for (int64_t j = 0; j < P1; ++j)
        for (int64_t i = 0; i < P0; ++i) {
            for (int64_t k = 0; k < n[2]; ++k) y[k] = at(i, j, k + 1);
            prefilter_line(y.data(), c.data(), n[2], cp, dp);
            for (int64_t k = 0; k < P2; ++k) at(i, j, k) = c[k];
        }
}

// nearest knot and offset along one axis: t = (x - x0)/dx + 1 (1-based data index),
// ix = round-half-up(t) limited to [1, n], δ = t - ix
inline void sdf_axis(const SDF& s, int a, real x, int64_t& ix, real& dl) {
    real t = (x - s.x0[a]) / s.d[a] + 1.0;
    real r = floor(t + 0.5);
    if (r < 1) r = 1;
    if (r > (real)s.n[a]) r = (real)s.n[a];
    ix = (int64_t)r;
    dl = t - r;
}

// value, gradient, Hessian of the interpolant (Interpolations.jl value/gradient/hessian weights)
inline void sdf_eval(const SDF& s, const V3& p, real& val, V3& grad, M3& hess, bool need_derivs) {
    int64_t ix[3]; real dl[3];
    real w[3][3], wd[3][3], wdd[3][3];
    for (int a = 0; a < 3; ++a) {
        sdf_axis(s, a, p[a], ix[a], dl[a]);
        real dd = dl[a];
        w[a][0] = (dd - 0.5) * (dd - 0.5) / 2; w[a][1] = 0.75 - dd * dd; w[a][2] = (dd + 0.5) * (dd + 0.5) / 2;
        wd[a][0] = (dd - 0.5) / s.d[a]; wd[a][1] = (-2 * dd) / s.d[a]; wd[a][2] = (dd + 0.5) / s.d[a];
        real h2 = s.d[a] * s.d[a];
        wdd[a][0] = 1.0 / h2; wdd[a][1] = -2.0 / h2; wdd[a][2] = 1.0 / h2;
    }
    // padded coefficient index of knot ix (1-based data index) is ix; neighbours ix-1, ix+1
    real v = 0, gx = 0, gy = 0, gz = 0, hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0, hyz = 0;
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
                real q = s.c(ix[0] - 1 + a, ix[1] - 1 + b, ix[2] - 1 + c);
                v += w[0][a] * w[1][b] * w[2][c] * q;
                if (need_derivs) {
                    gx += wd[0][a] * w[1][b] * w[2][c] * q;
                    gy += w[0][a] * wd[1][b] * w[2][c] * q;
                    gz += w[0][a] * w[1][b] * wd[2][c] * q;
                    hxx += wdd[0][a] * w[1][b] * w[2][c] * q;
                    hyy += w[0][a] * wdd[1][b] * w[2][c] * q;
                    hzz += w[0][a] * w[1][b] * wdd[2][c] * q;
                    hxy += wd[0][a] * wd[1][b] * w[2][c] * q;
                    hxz += wd[0][a] * w[1][b] * wd[2][c] * q;
                    hyz += w[0][a] * wd[1][b] * wd[2][c] * q;
                }
            }
    val = v;
    if (need_derivs) {
        grad = v3(gx, gy, gz);
        hess = rows(hxx, hxy, hxz, hxy, hyy, hyz, hxz, hyz, hzz);
    }
}

struct Surface {
    int kind = 0;  // EB_SURFACE_*
    V3 center; real R = 0;
    SDF sdf;
};

// isinside, utils_compute_interactions_beams.jl:39-47
inline bool isinside(const V3& p, const real* dom) {
    real lx = p[0] - dom[0]; bool fx = lx >= 0 && lx <= (dom[1] - dom[0]);
    real ly = p[1] - dom[2]; bool fy = ly >= 0 && ly <= (dom[3] - dom[2]);
    real lz = p[2] - dom[4]; bool fz = lz >= 0 && lz <= (dom[5] - dom[4]);
    return fx && fy && fz;
}

// incontact_beams, utils_compute_interactions_beams.jl:23-36 (sphere) and :67-75 (SDF)
inline bool incontact(const V3& p, const Surface& s, real rb) {
    real gbar = rb / 4;
    if (s.kind == 1) {
        V3 aux = p - s.center;
        real g = (norm(aux) - s.R) - rb;
        return g <= gbar;
    }
    if (!isinside(p, s.sdf.dom)) return false;
    real v; V3 gr; M3 h;
    sdf_eval(s.sdf, p, v, gr, h, false);
    return v - rb <= gbar;
}

// contact_gap_beams, utils_compute_interactions_beams.jl:2-20 (sphere) and :50-64 (SDF)
inline void contact_gap(const V3& p, const Surface& s, real rb, real& g, V3& n, M3& H) {
    if (s.kind == 1) {
        V3 aux = p - s.center;
        real na = norm(aux);
        g = (na - s.R) - rb;
        real inv = 1 / na;
        n = inv * aux;
        H = inv * eye() - ((inv * inv) * inv) * outer(aux, aux);
        return;
    }
    real v; V3 gr; M3 h;
    sdf_eval(s.sdf, p, v, gr, h, true);
    real nn = dot(gr, gr);
    real nmaginv = 1 / sqrt(nn);
    n = gr * nmaginv;
    H = (h * nmaginv) * (eye() - outer(n, n) / nn);
    g = v - rb;
}

// regularize_gap_penalty_beams, interactions_beams_regularization.jl:2-26
inline void regularize_gap_penalty(real g, real rb, real& p, real& dp, real& Pi) {
    real gb = rb / 4, h = gb / 2;
    if (g <= 0) {
        p = h - g; dp = -1.0;
        Pi = ((g * g) / 2 - h * g) + (gb * gb) / 6;
    } else {
        real aux = (gb - h) / (gb * gb);
        p = (aux * (g * g) - g) + h;
        dp = (2 * aux) * g - 1;
        Pi = (((gb - h) / (3 * (gb * gb))) * ((g * g) * g) - (g * g) / 2 + h * g) - (gb * gb) / 6;
    }
}

// smoothstep_transition_beams, interactions_beams_regularization.jl:29-45
inline void smoothstep(real tv, real pos, real rb, real& v, real& dv) {
    real ul = rb / 4;
    if (pos <= 0) { v = tv; dv = 0.0; return; }
    real s = (pos - ul) / ul;
    v = ((s * s) * (3 + 2 * s)) * tv;
    dv = (((6 / ul) * s) * (1 + s)) * tv;
}

struct ContactOut {
    real energy;
    real Tc[12];
    Mat12 K, C;
    int active;  // number of Gauss points in contact
};

// compute_contact_beams, src/interactions/compute&assemble_interactions_beams.jl:1-445
inline void compute_contact_beams(const ElemIn& in, real rb, const ContactProps& cp, const Surface& surf,
                                  const Gauss& g, ContactOut& out) {
    const M3 I = eye();
    V3 x1 = in.X1 + in.u1, x2 = in.X2 + in.u2;
    real ln = norm(x2 - x1);
    LocalFrame f = local_Re_and_aux(x1, x2, in.R1, in.R2, col(in.Re0, 1), ln);
    const M3& Re = f.Re;
    M3 ReT = tr(Re);
    const M3 &G1 = f.G1, &G2 = f.G2, &G3 = f.G3, &G4 = f.G4;
    M3 Rb1 = (ReT * in.R1) * in.Re0;
    M3 Rb2 = (ReT * in.R2) * in.Re0;
    V3 Tb1 = toangle(Rb1), Tb2 = toangle(Rb2);
    V3 Ud1 = ReT * in.ud1, Ud2 = ReT * in.ud2, Wd1 = ReT * in.wd1, Wd2 = ReT * in.wd2;
    M3 SUd1 = skew(Ud1), SUd2 = skew(Ud2), SWd1 = skew(Wd1), SWd2 = skew(Wd2);
    M3 G1R = G1 * ReT, G2R = G2 * ReT, G4R = G4 * ReT;
    M3 RG1 = Re * tr(G1), RG2 = Re * tr(G2), RG4 = Re * tr(G4);
    real energy = 0;
    M3 Kc[4][4], Cc[4][4];
    V3 Tc[4];
    for (int a = 0; a < 4; ++a) { Tc[a] = zv(); for (int b = 0; b < 4; ++b) { Kc[a][b] = zm(); Cc[a][b] = zm(); } }
    real eta_n = cp.eta_n;  // overwritten by the smoothstep at each contacting Gauss point (:100)
    int active = 0;
    for (int ig = 0; ig < g.n; ++ig) {
        real z = g.z[ig], w = g.w[ig];
        Shape s = shape(in.l0, z);
        V3 ut = v3(0, s.N3 * Tb1[2] + s.N4 * Tb2[2], (-s.N3) * Tb1[1] + (-s.N4) * Tb2[1]);
        M3 Sut = skew(ut);
        V3 xG = (s.N1 * x1 + s.N2 * x2) + Re * ut;
        real N7l = s.N7 / ln, N7l2 = N7l / ln;
        M3 PP1 = rows(0, 0, 0, 0, N7l, 0, 0, 0, N7l);
        M3 PP2 = rows(0, 0, 0, 0, 0, s.N3, 0, -s.N3, 0);
        M3 PP3 = -PP1;
        M3 PP4 = rows(0, 0, 0, 0, 0, s.N4, 0, -s.N4, 0);
        M3 H1[4];
        H1[0] = (s.N1 * I + PP1) - Sut * G1;
        H1[1] = PP2 - Sut * G2;
        H1[2] = (s.N2 * I + PP3) - Sut * G3;
        H1[3] = PP4 - Sut * G4;
        V3 h1 = ((H1[0] * Ud1 + H1[1] * Wd1) + H1[2] * Ud2) + H1[3] * Wd2;
        M3 Sh1 = skew(h1);
        M3 H1F1 = ((H1[0] * SUd1 + H1[1] * SWd1) + H1[2] * SUd2) + H1[3] * SWd2;
        V3 ud0 = Re * h1;
        if (!incontact(xG, surf, rb)) continue;
        ++active;
        real gn; V3 n; M3 Hg;
        contact_gap(xG, surf, rb, gn, n, Hg);
        real pn, dpn, Pi;
        regularize_gap_penalty(gn, rb, pn, dpn, Pi);
        real deta;
        smoothstep(eta_n, gn, rb, eta_n, deta);
        real udn_mag = dot(ud0, n);
        V3 udn = udn_mag * n;
        V3 udt = ud0 - udn;
        real udt2 = dot(udt, udt);
        energy += (w * cp.kn) * Pi;
        V3 fn = (cp.kn * pn) * n - eta_n * udn;
        real mureg = cp.mu / sqrt(udt2 + cp.eps_t);
        V3 ft = (((-cp.kn) * pn) * mureg) * udt;
        V3 fc = fn + ft;
        V3 Fc = ReT * fc;
        M3 RHR[4];
        for (int a = 0; a < 4; ++a) RHR[a] = (Re * H1[a]) * ReT;
        for (int a = 0; a < 4; ++a) Tc[a] += w * (tr(RHR[a]) * fc);
        M3 t1[4][4], t2[4][4], t3[4][4], t4[4][4];
        for (int a = 0; a < 4; ++a) {
            M3 RS = Re * skew(tr(H1[a]) * Fc);
            t1[a][0] = -(RS * G1R);
            t1[a][1] = -(RS * G2R);
            t1[a][2] = -t1[a][0];
            t1[a][3] = -(RS * G4R);
        }
        V3 v1 = f.r3;
        M3 A1F = rows(0, 0, 0, Fc[1] * v1[0], Fc[1] * v1[1], Fc[1] * v1[2], Fc[2] * v1[0], Fc[2] * v1[1], Fc[2] * v1[2]);
        M3 SF = skew(Fc);
        M3 SFP1 = (SF * PP1) * ReT, SFP2 = (SF * PP2) * ReT, SFP4 = (SF * PP4) * ReT;
        t2[0][0] = (N7l2 * Re) * A1F - RG1 * SFP1;
        t2[0][1] = -(RG1 * SFP2);
        t2[0][2] = -t2[0][0];
        t2[0][3] = -(RG1 * SFP4);
        t2[1][0] = -(RG2 * SFP1);
        t2[1][1] = -(RG2 * SFP2);
        t2[1][2] = -t2[1][0];
        t2[1][3] = -(RG2 * SFP4);
        t2[2][0] = t2[0][2];
        t2[2][1] = -t2[0][1];
        t2[2][2] = t2[0][0];
        t2[2][3] = -t2[0][3];
        t2[3][0] = -(RG4 * SFP1);
        t2[3][1] = -(RG4 * SFP2);
        t2[3][2] = -t2[3][0];
        t2[3][3] = -(RG4 * SFP4);
        for (int a = 0; a < 4; ++a) {
            M3 RHS = (Re * tr(H1[a])) * SF;
            t3[a][0] = RHS * G1R;
            t3[a][1] = RHS * G2R;
            t3[a][2] = -t3[a][0];
            t3[a][3] = RHS * G4R;
        }
        M3 aux = dot(n, ud0) * Hg + outer(n, Hg * ud0);
        M3 A1c[4];
        for (int a = 0; a < 4; ++a) A1c[a] = aux * RHR[a];
        M3 RSh1 = Re * Sh1;
        M3 A2_1 = -(RSh1 * G1R), A2_2 = -(RSh1 * G2R), A2_4 = -(RSh1 * G4R);
        M3 A1Dr = rows(0, 0, 0, v1[0] * (Ud1[1] - Ud2[1]), v1[1] * (Ud1[1] - Ud2[1]), v1[2] * (Ud1[1] - Ud2[1]),
                       v1[0] * (Ud1[2] - Ud2[2]), v1[1] * (Ud1[2] - Ud2[2]), v1[2] * (Ud1[2] - Ud2[2]));
        M3 RSGD = Re * skew(((G1 * Ud1 + G2 * Wd1) + G3 * Ud2) + G4 * Wd2);
        M3 A3_1 = (N7l2 * Re) * A1Dr + (RSGD * PP1) * ReT;
        M3 A3_2 = (RSGD * PP2) * ReT;
        M3 A3_4 = (RSGD * PP4) * ReT;
        M3 RH1SD = Re * H1F1;
        M3 A4_1 = RH1SD * G1R, A4_2 = RH1SD * G2R, A4_4 = RH1SD * G4R;
        M3 du0dd[4];
        du0dd[0] = (A2_1 + A3_1) + A4_1;
        du0dd[1] = (A2_2 + A3_2) + A4_2;
        du0dd[2] = -du0dd[0];
        du0dd[3] = (A2_4 + A3_4) + A4_4;
        M3 nn = outer(n, n);
        M3 dundd[4], dutdd[4], dundv[4], du0dv[4], dutdv[4];
        for (int a = 0; a < 4; ++a) {
            dundd[a] = A1c[a] + nn * du0dd[a];
            dutdd[a] = du0dd[a] - dundd[a];
        }
        for (int a = 0; a < 4; ++a) {
            dundv[a] = nn * RHR[a];
            du0dv[a] = RHR[a];
            dutdv[a] = du0dv[a] - dundv[a];
        }
        M3 auxn = ((cp.kn * dpn) * nn + (cp.kn * pn) * Hg) - outer(deta * udn, n);
        M3 Kfn[4], Cfn[4], Kft[4], Cft[4], Kfc[4], Cfc[4];
        for (int a = 0; a < 4; ++a) {
            Kfn[a] = auxn * RHR[a] - eta_n * dundd[a];
            Cfn[a] = ((-eta_n) * nn) * du0dv[a];
        }
        M3 aux1 = outer((((-mureg) * cp.kn) * dpn) * udt, n);
        M3 aux2 = (((-mureg) * cp.kn) * pn) * (I - outer((1 / (udt2 + cp.eps_t)) * udt, udt));
        for (int a = 0; a < 4; ++a) {
            Kft[a] = aux1 * RHR[a] + aux2 * dutdd[a];
            Cft[a] = aux2 * dutdv[a];
            Kfc[a] = Kfn[a] + Kft[a];
        }
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) t4[a][b] = tr(RHR[a]) * Kfc[b];
        for (int a = 0; a < 4; ++a) Cfc[a] = Cfn[a] + Cft[a];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                Kc[a][b] += w * (((t1[a][b] + t2[a][b]) + t3[a][b]) + t4[a][b]);
                Cc[a][b] += (w * tr(RHR[a])) * Cfc[b];
            }
    }
    real l02 = in.l0 / 2;
    for (int a = 0; a < 4; ++a) {
        Tc[a] = l02 * Tc[a];
        for (int b = 0; b < 4; ++b) { Cc[a][b] = l02 * Cc[a][b]; Kc[a][b] = l02 * Kc[a][b]; }
    }
    out.energy = l02 * energy;
    V3 Th1 = toangle(in.dR1), Th2 = toangle(in.dR2);
    M3 TT1t = tr(Ts_inv(Th1)), TT2t = tr(Ts_inv(Th2));
    for (int b = 0; b < 4; ++b) { Cc[1][b] = Cc[1][b] * TT1t; Cc[3][b] = Cc[3][b] * TT2t; }
    for (int a = 0; a < 4; ++a) set_vec(out.Tc, a, Tc[a]);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) { out.K.set_block(a, b, Kc[a][b]); out.C.set_block(a, b, Cc[a][b]); }
    out.active = active;
}

}  // namespace ebo
