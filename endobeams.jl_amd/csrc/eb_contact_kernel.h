// eb_contact_kernel.h — beam-to-rigid penalty/friction contact at Gauss points
// (compute_contact_beams, src/interactions/compute&assemble_interactions_beams.jl:1-445) with the
// gap functions of src/interactions/utils_compute_interactions_beams.jl:1-75 and the
// regularisation of src/interactions/interactions_beams_regularization.jl:1-45.
//
// Two passes: a detection pass (one lane per contact beam) places the Gauss points (xᴳ), runs the
// contact test and lists the beams in contact; the tangent pass spreads each listed beam over four
// wavefronts, one per block column of Kᶜ/Cᶜ. Contributions are
// folded into the element staging in place: K -= Kᶜ, C -= (1+α)Cᶜ (assemble_contact!, :493-497),
// so one gather assembles beams and contact together.
#pragma once
#include "eb_beam_kernel.h"

namespace ebd {

// Quadratic B-spline SDF with Reflect(OnCell()) prefiltered coefficients, padded by one cell
// (src/interactions/def_interactions.jl:67-82).
// Coarse lower bound of the interpolant: cell c of axis a holds the stencil centres (1-based
// knots) B·c + 1 … B·c + B, and its entry the minimum of the coefficients any of their stencils
// reads (padded indices B·c … B·c + B + 1), so that φ(x) ≥ lower bound for every x whose stencil
// centre falls in the cell (the quadratic B-spline weights are non-negative and sum to one).
constexpr int SDF_BLOCK = 8;

template <typename P> struct SDFDev {
    const P* coef;       // (n0+2)(n1+2)(n2+2), x fastest
    int32_t n[3];
    double d[3], x0[3];
    double dom[6];
    const P* cmin;       // coarse lower bounds, cn0·cn1·cn2, x fastest
    int32_t cn[3];       // ceil(n / SDF_BLOCK)
};

// stencil centre (1-based knot) of coordinate x along an axis, as sdf_eval computes it
EBD int sdf_knot(double x, double x0, double d, int32_t n) {
    double t = (x - x0) / d + 1.0;
    double r = floor(t + 0.5);
    r = r < 1.0 ? 1.0 : r;
    r = r > double(n) ? double(n) : r;
    return int(r);
}

struct SphereDev { double c[3], R; };

template <typename P> struct ContactArgs {
    int surface;  // EB_SURFACE_*
    const SphereDev* sphere;   // [ninst]
    const SDFDev<P>* sdf;      // [ninst]
    const int32_t* beam_inst;  // [E]
    const uint8_t* active;     // [ninst]
    double kn, mu, eps_t, eta_n, alpha;
    const int32_t* beams;  // 0-based contact beam list
    int64_t n;
    double* Tc;            // [E][12] staging (written for every listed beam), FP64
    double* energy;        // [E][3] staging, contact energy in slot 2
    unsigned long long* stats;   // [2]: listed beams and their Gauss points in contact, summed over launches
};

EBD bool isinside_dom(V3<double> p, const double* dom) {
    double lx = p.x - dom[0], ly = p.y - dom[2], lz = p.z - dom[4];
    return lx >= 0 && lx <= (dom[1] - dom[0]) && ly >= 0 && ly <= (dom[3] - dom[2]) && lz >= 0 && lz <= (dom[5] - dom[4]);
}

// 27-coefficient stencil: value (and optionally gradient, Hessian) of the interpolant.
// t = (x - x₀)/dx + 1, knot = round-half-up(t) clamped to [1, n], δ = t - knot.
template <typename P, typename C>
EBD void sdf_eval(const SDFDev<C>& s, V3<double> p, P& val, V3<P>* grad, M3<P>* hess) {
    int ix[3];
    P w[3][3], wd[3][3], wdd[3][3];
    double pc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t = (pc[a] - s.x0[a]) / s.d[a] + 1.0;
        double r = floor(t + 0.5);
        r = r < 1.0 ? 1.0 : r;
        r = r > double(s.n[a]) ? double(s.n[a]) : r;
        ix[a] = int(r);
        P dd = P(t - r);
        P h = P(0.5);
        w[a][0] = (dd - h) * (dd - h) / P(2); w[a][1] = P(0.75) - dd * dd; w[a][2] = (dd + h) * (dd + h) / P(2);
        P id = P(1.0 / s.d[a]);
        wd[a][0] = (dd - h) * id; wd[a][1] = (P(-2) * dd) * id; wd[a][2] = (dd + h) * id;
        P ih2 = P(1.0 / (s.d[a] * s.d[a]));
        wdd[a][0] = ih2; wdd[a][1] = P(-2) * ih2; wdd[a][2] = ih2;
    }
    const int64_t P0 = s.n[0] + 2, P1 = s.n[1] + 2;
    P v = 0, gx = 0, gy = 0, gz = 0, hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0, hyz = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const C* row = s.coef + ((int64_t)(ix[2] - 1 + c) * P1 + (ix[1] - 1 + b)) * P0 + (ix[0] - 1);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                P q = P(row[a]);
                v += w[0][a] * w[1][b] * w[2][c] * q;
                if (grad) {
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
        }
    val = v;
    if (grad) {
        *grad = V3<P>{gx, gy, gz};
        *hess = rows(hxx, hxy, hxz, hxy, hyy, hyz, hxz, hyz, hzz);
    }
}

// incontact_beams (utils_compute_interactions_beams.jl:23-36, :67-75)
template <typename P> EBD bool in_contact(const ContactArgs<P>& ca, int32_t inst, V3<double> p, double rb) {
    double gbar = rb / 4;
    if (ca.surface == 1) {
        const SphereDev& sp = ca.sphere[inst];
        V3<double> aux = {p.x - sp.c[0], p.y - sp.c[1], p.z - sp.c[2]};
        return (norm(aux) - sp.R) - rb <= gbar;
    }
    const SDFDev<P>& sdf = ca.sdf[inst];
    if (!isinside_dom(p, sdf.dom)) return false;
    // far from the surface: the coarse lower bound already exceeds the contact threshold (with a
    // margin far above the rounding of the 27-term sum), so the full stencil would say the same
    const int c0 = (sdf_knot(p.x, sdf.x0[0], sdf.d[0], sdf.n[0]) - 1) / SDF_BLOCK;
    const int c1 = (sdf_knot(p.y, sdf.x0[1], sdf.d[1], sdf.n[1]) - 1) / SDF_BLOCK;
    const int c2 = (sdf_knot(p.z, sdf.x0[2], sdf.d[2], sdf.n[2]) - 1) / SDF_BLOCK;
    const double lb = double(sdf.cmin[((int64_t)c2 * sdf.cn[1] + c1) * sdf.cn[0] + c0]);
    if (lb > 0 && lb * (1 - 1e-12) - rb > gbar + 1e-12 * (lb + rb)) return false;
    double v;
    sdf_eval<double, P>(sdf, p, v, nullptr, nullptr);
    return v - rb <= gbar;
}

// contact_gap_beams (utils_compute_interactions_beams.jl:2-20, :50-64)
template <typename P, typename C>
EBD void contact_gap(const ContactArgs<C>& ca, int32_t inst, V3<double> p, double rb, P& g, V3<P>& n, M3<P>& H) {
    if (ca.surface == 1) {
        const SphereDev& sp = ca.sphere[inst];
        V3<P> aux = {P(p.x - sp.c[0]), P(p.y - sp.c[1]), P(p.z - sp.c[2])};
        P na = norm(aux);
        g = (na - P(sp.R)) - P(rb);
        P inv = P(1) / na;
        n = inv * aux;
        H = inv * eye<P>() - ((inv * inv) * inv) * outer(aux, aux);
        return;
    }
    P v;
    V3<P> gr;
    M3<P> h;
    sdf_eval<P, C>(ca.sdf[inst], p, v, &gr, &h);
    P nn = dot(gr, gr);
    P nmaginv = P(1) / sqrt(nn);
    n = gr * nmaginv;
    H = (h * nmaginv) * (eye<P>() - outer(n, n) / nn);
    g = v - P(rb);
}

// Consistent contact tangent of one beam for block column B of Kᶜ and Cᶜ (rows a = 0..3), and
// Tᶜ of node slot B; the beam has a Gauss point in contact (mask: the points in contact).
// B = CONTACT_ALL_SLOTS (force only): Tᶜ of all four slots and the contact energy in one pass.
enum { CONTACT_FORCE = 1, CONTACT_TANGENT = 2 };
constexpr int CONTACT_ALL_SLOTS = 4;

template <typename P, typename C, int B, int WHAT>
EBD void contact_column(int64_t e, unsigned mask, const NodeState& ns, const BeamConsts& bc, const BeamStage<C>& st,
                        const GaussPts& gp, const ContactArgs<C>& ca) {
    constexpr bool FORCE = (WHAT & CONTACT_FORCE) != 0, TANGENT = (WHAT & CONTACT_TANGENT) != 0;
    constexpr bool ALL = B == CONTACT_ALL_SLOTS;
    static_assert(!(ALL && TANGENT), "the tangent is computed one block column at a time");
    constexpr int BC = ALL ? 0 : B;   // the block column of the tangent
    const int64_t E = bc.E;
    const int32_t inst = ca.beam_inst[e];
    const int32_t na = bc.n1[e], nb = bc.n2[e];
    const double rb = bc.props[0 * E + e];
    V3<double> x1 = ldv(ns.X0, na) + ldv(ns.u, na), x2 = ldv(ns.X0, nb) + ldv(ns.u, nb);
    double lnd = norm(x2 - x1);
    M3<double> Re0;
    for (int k = 0; k < 9; ++k) Re0.a[k] = bc.Re0[k * E + e];
    M3<double> R1 = ldm(ns.R, na), R2 = ldm(ns.R, nb);
    LocalFrame f = local_frame(x1, x2, R1, R2, col(Re0, 1), lnd);
    M3<double> ReTd = tr(f.Re);
    V3<double> Tb1 = toangle((ReTd * R1) * Re0), Tb2 = toangle((ReTd * R2) * Re0);
    const double l0 = bc.l0[e];

    const M3<P> I = eye<P>();
    const M3<P> Re = cvt<P>(f.Re), ReT = tr(Re);
    const P ln = P(lnd);
    const M3<P> G1 = rows(P(0), P(0), P(f.eta / lnd), P(0), P(0), P(1 / lnd), P(0), P(-1 / lnd), P(0));
    const M3<P> G2 = rows(P(f.eta12 / 2), P(-f.eta11 / 2), P(0), P(0), P(0), P(0), P(0), P(0), P(0));
    const M3<P> G3 = -G1;
    const M3<P> G4 = rows(P(f.eta22 / 2), P(-f.eta21 / 2), P(0), P(0), P(0), P(0), P(0), P(0), P(0));
    const V3<P> Ud1 = tmul(Re, cvt<P>(ldv(ns.ud, na))), Ud2 = tmul(Re, cvt<P>(ldv(ns.ud, nb)));
    const V3<P> Wd1 = tmul(Re, cvt<P>(ldv(ns.wd, na))), Wd2 = tmul(Re, cvt<P>(ldv(ns.wd, nb)));
    const M3<P> SUd1 = skew(Ud1), SUd2 = skew(Ud2), SWd1 = skew(Wd1), SWd2 = skew(Wd2);
    const M3<P> G1R = G1 * ReT, G2R = G2 * ReT, G4R = G4 * ReT;
    const M3<P> RG1 = Re * tr(G1), RG2 = Re * tr(G2), RG4 = Re * tr(G4);
    const V3<P> v1 = cvt<P>(f.v1);
    const P kn = P(ca.kn), mu = P(ca.mu), eps_t = P(ca.eps_t);
    const V3<P> tb1 = cvt<P>(Tb1), tb2 = cvt<P>(Tb2);
    P eta_n = P(ca.eta_n);  // overwritten by the smoothstep at each contacting point (:100)
    double energy = 0;
    V3<P> Tcs[ALL ? 4 : 1];
    for (auto& t : Tcs) t = zv<P>();
    M3<P> Kc[4], Cc[4];
    for (int a = 0; a < 4; ++a) { Kc[a] = zm<P>(); Cc[a] = zm<P>(); }
#pragma unroll 1
    for (int ig = 0; ig < gp.n; ++ig) {
        if (!(mask & (1u << ig))) continue;
        const P w = P(gp.w[ig]);
        ShapeF sd = shape_fn(l0, gp.z[ig]);
        V3<double> utd = {0.0, sd.N3 * Tb1.z + sd.N4 * Tb2.z, (-sd.N3) * Tb1.y + (-sd.N4) * Tb2.y};
        V3<double> xG = (sd.N1 * x1 + sd.N2 * x2) + f.Re * utd;
        const P N1 = P(sd.N1), N2 = P(sd.N2), N3 = P(sd.N3), N4 = P(sd.N4), N7 = P(sd.N7);
        V3<P> ut = {P(0), N3 * tb1.z + N4 * tb2.z, (-N3) * tb1.y + (-N4) * tb2.y};
        M3<P> Sut = skew(ut);
        P N7l = N7 / ln, N7l2 = N7l / ln;
        M3<P> PP1 = rows(P(0), P(0), P(0), P(0), N7l, P(0), P(0), P(0), N7l);
        M3<P> PP2 = rows(P(0), P(0), P(0), P(0), P(0), N3, P(0), -N3, P(0));
        M3<P> PP3 = -PP1;
        M3<P> PP4 = rows(P(0), P(0), P(0), P(0), P(0), N4, P(0), -N4, P(0));
        M3<P> H1[4];
        H1[0] = (N1 * I + PP1) - Sut * G1;
        H1[1] = PP2 - Sut * G2;
        H1[2] = (N2 * I + PP3) - Sut * G3;
        H1[3] = PP4 - Sut * G4;
        V3<P> h1 = ((H1[0] * Ud1 + H1[1] * Wd1) + H1[2] * Ud2) + H1[3] * Wd2;
        M3<P> Sh1 = skew(h1);
        M3<P> H1F1 = ((H1[0] * SUd1 + H1[1] * SWd1) + H1[2] * SUd2) + H1[3] * SWd2;
        V3<P> ud0 = Re * h1;

        P gn;
        V3<P> n;
        M3<P> Hg;
        contact_gap<P, C>(ca, inst, xG, rb, gn, n, Hg);
        // regularize_gap_penalty_beams
        P gb = P(rb) / P(4), hh = gb / P(2), pn, dpn, Pi;
        if (gn <= P(0)) {
            pn = hh - gn; dpn = P(-1);
            Pi = ((gn * gn) / P(2) - hh * gn) + (gb * gb) / P(6);
        } else {
            P aux = (gb - hh) / (gb * gb);
            pn = (aux * (gn * gn) - gn) + hh;
            dpn = (P(2) * aux) * gn - P(1);
            Pi = (((gb - hh) / (P(3) * (gb * gb))) * ((gn * gn) * gn) - (gn * gn) / P(2) + hh * gn) - (gb * gb) / P(6);
        }
        // smoothstep_transition_beams
        P deta;
        if (gn <= P(0)) {
            deta = P(0);
        } else {
            P ss = (gn - gb) / gb;
            deta = (((P(6) / gb) * ss) * (P(1) + ss)) * eta_n;
            eta_n = ((ss * ss) * (P(3) + P(2) * ss)) * eta_n;
        }
        P udn_mag = dot(ud0, n);
        V3<P> udn = udn_mag * n;
        V3<P> udt = ud0 - udn;
        P udt2 = dot(udt, udt);
        if (FORCE) energy += double((w * kn) * Pi);
        V3<P> fn = (kn * pn) * n - eta_n * udn;
        P mureg = mu / sqrt(udt2 + eps_t);
        V3<P> ft = (((-kn) * pn) * mureg) * udt;
        V3<P> fc = fn + ft;
        V3<P> Fc = tmul(Re, fc);
        M3<P> RHR[4];
        for (int a = 0; a < 4; ++a) RHR[a] = (Re * H1[a]) * ReT;
        if constexpr (FORCE) {
            if constexpr (ALL) {
                for (int a = 0; a < 4; ++a) Tcs[a] += w * tmul(RHR[a], fc);
            } else {
                Tcs[0] += w * tmul(RHR[BC], fc);
            }
        }
        if (!TANGENT) continue;
        M3<P> A1F = rows(P(0), P(0), P(0), Fc.y * v1.x, Fc.y * v1.y, Fc.y * v1.z, Fc.z * v1.x, Fc.z * v1.y, Fc.z * v1.z);
        M3<P> SF = skew(Fc);
        M3<P> SFP1 = (SF * PP1) * ReT, SFP2 = (SF * PP2) * ReT, SFP4 = (SF * PP4) * ReT;
        M3<P> aux = dot(n, ud0) * Hg + outer(n, Hg * ud0);
        M3<P> RSh1 = Re * Sh1;
        M3<P> A1Dr = rows(P(0), P(0), P(0), v1.x * (Ud1.y - Ud2.y), v1.y * (Ud1.y - Ud2.y), v1.z * (Ud1.y - Ud2.y),
                          v1.x * (Ud1.z - Ud2.z), v1.y * (Ud1.z - Ud2.z), v1.z * (Ud1.z - Ud2.z));
        M3<P> RSGD = Re * skew(((G1 * Ud1 + G2 * Wd1) + G3 * Ud2) + G4 * Wd2);
        M3<P> RH1SD = Re * H1F1;
        M3<P> du0dd[4];
        du0dd[0] = ((-(RSh1 * G1R)) + ((N7l2 * Re) * A1Dr + (RSGD * PP1) * ReT)) + RH1SD * G1R;
        du0dd[1] = ((-(RSh1 * G2R)) + (RSGD * PP2) * ReT) + RH1SD * G2R;
        du0dd[2] = -du0dd[0];
        du0dd[3] = ((-(RSh1 * G4R)) + (RSGD * PP4) * ReT) + RH1SD * G4R;
        M3<P> nn = outer(n, n);
        M3<P> auxn = ((kn * dpn) * nn + (kn * pn) * Hg) - outer(deta * udn, n);
        M3<P> aux1 = outer((((-mureg) * kn) * dpn) * udt, n);
        M3<P> aux2 = (((-mureg) * kn) * pn) * (I - outer((P(1) / (udt2 + eps_t)) * udt, udt));
        constexpr int b = BC;
        M3<P> dundd = aux * RHR[b] + nn * du0dd[b];
        M3<P> dutdd = du0dd[b] - dundd;
        M3<P> dutdv = RHR[b] - nn * RHR[b];
        M3<P> Kfn = auxn * RHR[b] - eta_n * dundd;
        M3<P> Cfn = ((-eta_n) * nn) * RHR[b];
        M3<P> Kft = aux1 * RHR[b] + aux2 * dutdd;
        M3<P> Cft = aux2 * dutdv;
        const M3<P> Kfc = Kfn + Kft;
        const M3<P> Cfc = Cfn + Cft;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            M3<P> RS = Re * skew(tmul(H1[a], Fc));
            M3<P> RHS = (Re * tr(H1[a])) * SF;
            M3<P> tRHR = tr(RHR[a]);
            M3<P> RGa = a == 0 ? RG1 : (a == 1 ? RG2 : (a == 2 ? -RG1 : RG4));
            M3<P> t1, t2, t3;
            // t₁ᵃᵇ = -Rₑ S(H₁ᵃᵀ𝓕ᶜ) Gᵇ Rₑᵀ with G³ = -G¹
            M3<P> GbR = b == 0 ? G1R : (b == 1 ? G2R : (b == 2 ? G1R : G4R));
            t1 = b == 2 ? RS * GbR : -(RS * GbR);
            t3 = b == 2 ? -(RHS * GbR) : RHS * GbR;
            // t₂ (:166-184): rows 1 and 3 carry the A₁ᵀ𝓕ᶜr term; row 3 mirrors row 1
            M3<P> SFPb = b == 0 ? SFP1 : (b == 1 ? SFP2 : (b == 2 ? SFP1 : SFP4));
            if (a == 0 || a == 2) {
                M3<P> base;
                if (b == 0 || b == 2) base = (N7l2 * Re) * A1F - RG1 * SFP1;
                else base = -(RG1 * SFPb);
                // t₂¹¹ = base₁, t₂¹³ = -base₁, t₂³¹ = t₂¹³, t₂³³ = t₂¹¹; t₂¹² = base₂, t₂³² = -t₂¹²; likewise col 4
                bool neg = (b == 2) != (a == 2);
                t2 = neg ? -base : base;
            } else {
                M3<P> base = -(RGa * SFP1);
                if (b == 0) t2 = base;
                else if (b == 2) t2 = -base;
                else t2 = -(RGa * SFPb);
            }
            M3<P> t4 = tRHR * Kfc;
            Kc[a] += w * (((t1 + t2) + t3) + t4);
            Cc[a] += (w * tRHR) * Cfc;
        }
    }
    const P l02 = P(l0 / 2);
    if (FORCE) {
        if (B == 0 || ALL) ca.energy[3 * e + 2] = double(l02) * energy;
        for (int k = 0; k < (ALL ? 4 : 1); ++k) {
            const int slot = ALL ? k : BC;
            V3<P> t = l02 * Tcs[k];
            ca.Tc[e * 12 + 3 * slot] = double(t.x);
            ca.Tc[e * 12 + 3 * slot + 1] = double(t.y);
            ca.Tc[e * 12 + 3 * slot + 2] = double(t.z);
        }
    }
    if (!TANGENT) return;
    const M3<P> TT1t = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, na)))));
    const M3<P> TT2t = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, nb)))));
    const P c1a = P(1 + ca.alpha);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        M3<P> K = l02 * Kc[a];
        M3<P> Cm = l02 * Cc[a];
        if (a == 1) Cm = Cm * TT1t;
        if (a == 3) Cm = Cm * TT2t;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int64_t idx = stage_idx(e, 3 * a + i, 3 * BC + j);
                st.K[idx] += C(-K.a[i + 3 * j]);
                st.C[idx] += C(-(c1a * Cm.a[i + 3 * j]));
            }
    }
}

// Pass 1, one lane per listed contact beam: Gauss point placement (FP64 positions) and the
// contact test. Beams with no point in contact get zero Tᶜ and contact energy; the others are
// appended to the work list as (beam << 8 | point mask). List order does not matter: every beam
// writes only its own staging record.
template <typename P>
__global__ void __launch_bounds__(256) contact_detect(NodeState ns, BeamConsts bc, GaussPts gp, ContactArgs<P> ca,
                                                      uint64_t* list, int32_t* count) {
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= ca.n) return;
    const int64_t E = bc.E;
    const int64_t e = ca.beams[li];
    const int32_t inst = ca.beam_inst[e];
    if (!ca.active[inst]) return;   // frozen instance: its staging is not gathered
    const int32_t na = bc.n1[e], nb = bc.n2[e];
    const double rb = bc.props[0 * E + e];
    V3<double> x1 = ldv(ns.X0, na) + ldv(ns.u, na), x2 = ldv(ns.X0, nb) + ldv(ns.u, nb);
    double lnd = norm(x2 - x1);
    M3<double> Re0;
    for (int k = 0; k < 9; ++k) Re0.a[k] = bc.Re0[k * E + e];
    M3<double> R1 = ldm(ns.R, na), R2 = ldm(ns.R, nb);
    LocalFrame f = local_frame(x1, x2, R1, R2, col(Re0, 1), lnd);
    M3<double> ReTd = tr(f.Re);
    V3<double> Tb1 = toangle((ReTd * R1) * Re0), Tb2 = toangle((ReTd * R2) * Re0);
    const double l0 = bc.l0[e];
    unsigned mask = 0;
    for (int ig = 0; ig < gp.n; ++ig) {
        ShapeF s = shape_fn(l0, gp.z[ig]);
        V3<double> ut = {0.0, s.N3 * Tb1.z + s.N4 * Tb2.z, (-s.N3) * Tb1.y + (-s.N4) * Tb2.y};
        V3<double> xG = (s.N1 * x1 + s.N2 * x2) + f.Re * ut;
        if (in_contact<P>(ca, inst, xG, rb)) mask |= 1u << ig;
    }
    if (!mask) {
        for (int k = 0; k < 12; ++k) ca.Tc[e * 12 + k] = 0.0;
        ca.energy[3 * e + 2] = 0.0;
        return;
    }
    const int32_t slot = atomicAdd(count, 1);
    list[slot] = (uint64_t(e) << 8) | mask;
}

// The same detection as a hook of the element kernel: each of the two parts that run it tests its
// Gauss points with the element's frame (x₁, x₂, Rₑ, Θ̅₁, Θ̅₂), and once the tile is done the mass
// part appends the beams with points in contact to the work list, so the assembly with contact
// needs no separate detection pass.
template <typename P> struct DetectHook {
    static constexpr bool SPLIT = true;
    const ContactArgs<P>& ca;   // the kernel's arguments, by reference (no local copies)
    const uint8_t* is_contact;  // [E]
    uint64_t* list;
    int32_t* count;
    EBD unsigned test(int64_t e, const BeamConsts& bc, const GaussPts& gp, unsigned gsel, const V3<double>& x1,
                      const V3<double>& x2, const M3<double>& Re, const V3<double>& Tb1, const V3<double>& Tb2) const {
        if (!is_contact[e] || !gsel) return 0;
        const int64_t E = bc.E;
        const int32_t inst = ca.beam_inst[e];
        const double rb = bc.props[0 * E + e];
        const double l0 = bc.l0[e];
        unsigned mask = 0;
        for (int ig = 0; ig < gp.n; ++ig) {
            if (!((gsel >> ig) & 1)) continue;
            ShapeF s = shape_fn(l0, gp.z[ig]);
            V3<double> ut = {0.0, s.N3 * Tb1.z + s.N4 * Tb2.z, (-s.N3) * Tb1.y + (-s.N4) * Tb2.y};
            V3<double> xG = (s.N1 * x1 + s.N2 * x2) + Re * ut;
            if (in_contact<P>(ca, inst, xG, rb)) mask |= 1u << ig;
        }
        return mask;
    }
    EBD void finish(int64_t e, unsigned mask) const {
        if (!is_contact[e]) return;
        if (!mask) {
            for (int k = 0; k < 12; ++k) ca.Tc[e * 12 + k] = 0.0;
            ca.energy[3 * e + 2] = 0.0;
            return;
        }
        const int32_t slot = atomicAdd(count, 1);
        list[slot] = (uint64_t(e) << 8) | mask;
    }
};

template <typename S>
__global__ void __launch_bounds__(4 * BEAM_TILE) beam_detect_kernel(NodeState ns, BeamConsts bc, BeamStage<S> st,
                                                                     GaussPts gp, const uint8_t* __restrict__ active,
                                                                     const int32_t* __restrict__ beam_inst,
                                                                     ContactArgs<S> ca, const uint8_t* is_contact,
                                                                     uint64_t* list, int32_t* count) {
    beam_kernel_body<S, 0, 4>(blockIdx.x, ns, bc, st, gp, active, beam_inst, DetectHook<S>{ca, is_contact, list, count});
}
// the static and mass parts with the contact detection (the precision form's first launch; the
// Coriolis parts follow in beam_kernel_half<S, PART_C13>)
template <typename S>
__global__ void __launch_bounds__(2 * BEAM_TILE) beam_detect_kernel_half(NodeState ns, BeamConsts bc, BeamStage<S> st,
                                                                          GaussPts gp, const uint8_t* __restrict__ active,
                                                                          const int32_t* __restrict__ beam_inst,
                                                                          ContactArgs<S> ca, const uint8_t* is_contact,
                                                                          uint64_t* list, int32_t* count) {
    beam_kernel_body<S, 0, 2>(blockIdx.x, ns, bc, st, gp, active, beam_inst, DetectHook<S>{ca, is_contact, list, count});
}

// Precision form: the FP64 contact force (Tᶜ of all four node slots) and energy of every listed
// beam, one lane per beam, with the arithmetic of the FP64 pass of the tangent kernel.
template <typename P>
EBD void contact_force_item(int64_t li, int32_t nlisted, const NodeState& ns, const BeamConsts& bc, const BeamStage<P>& st,
                            const GaussPts& gp, const ContactArgs<P>& ca, const uint64_t* __restrict__ list) {
    if (li >= nlisted) return;
    const uint64_t ent = list[li];
    contact_column<double, P, CONTACT_ALL_SLOTS, CONTACT_FORCE>((int64_t)(ent >> 8), (unsigned)(ent & 0xffu), ns, bc, st,
                                                                gp, ca);
}
template <typename P>
__global__ void __launch_bounds__(256) contact_force_kernel(NodeState ns, BeamConsts bc, BeamStage<P> st, GaussPts gp,
                                                            ContactArgs<P> ca, const uint64_t* __restrict__ list,
                                                            const int32_t* __restrict__ count) {
    contact_force_item<P>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, *count, ns, bc, st, gp, ca, list);
}

// ---- the contact tangent in two phases --------------------------------------------------------
//
// Per beam and Gauss point, everything up to the derivatives of the contact force, 𝒦ᶠᶜᵇ = ∂𝓯ᶜ/∂dᵇ and
// 𝒞ᶠᶜᵇ = ∂𝓯ᶜ/∂ḋᵇ (b = 1..4), is shared by the sixteen blocks of Kᶜ and Cᶜ; block (a, b) combines
// them with quantities of row a alone (H₁ᵃ, RₑH₁ᵃRₑᵀ, Ŝ(H₁ᵃᵀ𝓕ᶜ), RₑGᵃ) and of the Gauss point
// (𝓕ᶜ, the P₁Pᵇ blocks). Phase 1: wavefront g of a 64-beam tile evaluates Gauss point g of each
// beam up to 𝒦ᶠᶜ, 𝒞ᶠᶜ and 𝓕ᶜ and leaves them in LDS (component-major, one column per beam, so a
// wavefront's accesses are consecutive); phase 2: wavefront a sums block row a over the Gauss
// points in their order. Every expression is the one of contact_column, evaluated once instead of
// once per block column.
constexpr int CONTACT_TILE = 64;
constexpr int CONTACT_GP_CHUNK = 3;   // Gauss points per beam at most (phase-1 wavefronts 0..2)

template <typename P, bool FORCE> struct ContactShared {
    P kfc[CONTACT_GP_CHUNK][36][CONTACT_TILE];   // 𝒦ᶠᶜᵇ, b-major, each 3x3 column-major
    P cfc[CONTACT_GP_CHUNK][36][CONTACT_TILE];   // 𝒞ᶠᶜᵇ
    P fc[CONTACT_GP_CHUNK][3][CONTACT_TILE];     // 𝓕ᶜ = Rₑᵀ𝓯ᶜ
    P gn[CONTACT_GP_CHUNK][CONTACT_TILE];        // gap (the smoothstep chain of ηₙ runs over the points)
    P tt[2][9][CONTACT_TILE];                    // Tₛ⁻¹(Θ₁)ᵀ, Tₛ⁻¹(Θ₂)ᵀ (C rows 2 and 4)
};
template <typename P> struct ContactShared<P, true> {
    P kfc[CONTACT_GP_CHUNK][36][CONTACT_TILE];
    P cfc[CONTACT_GP_CHUNK][36][CONTACT_TILE];
    P fc[CONTACT_GP_CHUNK][3][CONTACT_TILE];
    P gn[CONTACT_GP_CHUNK][CONTACT_TILE];
    P tt[2][9][CONTACT_TILE];
    double tc[CONTACT_GP_CHUNK][12][CONTACT_TILE];   // ωᴳ RₑH₁ᵃRₑᵀᵀ𝓯ᶜ, a-major
    double en[CONTACT_GP_CHUNK][CONTACT_TILE];       // ωᴳ kₙ Πₑ
};

// A listed beam's frame and the per-beam quantities both phases use, in the physics type P.
template <typename P> struct ContactBeam {
    int64_t e;
    unsigned mask;
    int32_t inst, na, nb;
    double rb, l0;
    V3<double> x1, x2, Tb1, Tb2;
    M3<double> Red;
    M3<P> Re, ReT, G1, G2, G4;
    V3<P> v1, tb1, tb2;
    P ln;
};

template <typename P, typename C>
EBD ContactBeam<P> contact_beam(int64_t e, unsigned mask, const NodeState& ns, const BeamConsts& bc, const ContactArgs<C>& ca) {
    ContactBeam<P> cb;
    const int64_t E = bc.E;
    cb.e = e;
    cb.mask = mask;
    cb.inst = ca.beam_inst[e];
    cb.na = bc.n1[e];
    cb.nb = bc.n2[e];
    cb.rb = bc.props[0 * E + e];
    cb.x1 = ldv(ns.X0, cb.na) + ldv(ns.u, cb.na);
    cb.x2 = ldv(ns.X0, cb.nb) + ldv(ns.u, cb.nb);
    const double lnd = norm(cb.x2 - cb.x1);
    M3<double> Re0;
    for (int k = 0; k < 9; ++k) Re0.a[k] = bc.Re0[k * E + e];
    const M3<double> R1 = ldm(ns.R, cb.na), R2 = ldm(ns.R, cb.nb);
    const LocalFrame f = local_frame(cb.x1, cb.x2, R1, R2, col(Re0, 1), lnd);
    const M3<double> ReTd = tr(f.Re);
    cb.Tb1 = toangle((ReTd * R1) * Re0);
    cb.Tb2 = toangle((ReTd * R2) * Re0);
    cb.Red = f.Re;
    cb.l0 = bc.l0[e];
    cb.Re = cvt<P>(f.Re);
    cb.ReT = tr(cb.Re);
    cb.ln = P(lnd);
    cb.G1 = rows(P(0), P(0), P(f.eta / lnd), P(0), P(0), P(1 / lnd), P(0), P(-1 / lnd), P(0));
    cb.G2 = rows(P(f.eta12 / 2), P(-f.eta11 / 2), P(0), P(0), P(0), P(0), P(0), P(0), P(0));
    cb.G4 = rows(P(f.eta22 / 2), P(-f.eta21 / 2), P(0), P(0), P(0), P(0), P(0), P(0), P(0));
    cb.v1 = cvt<P>(f.v1);
    cb.tb1 = cvt<P>(cb.Tb1);
    cb.tb2 = cvt<P>(cb.Tb2);
    return cb;
}

// The shape-function quantities of one Gauss point (N₁…N₄, N₇/lₙ, N₇/lₙ², uᵗ, the P₁P blocks).
template <typename P> struct GaussShape {
    P N1, N2, N7l, N7l2;
    M3<P> Sut, PP1, PP2, PP4;
};
template <typename P> EBD GaussShape<P> gauss_shape(const ContactBeam<P>& cb, const ShapeF& sd) {
    GaussShape<P> gs;
    gs.N1 = P(sd.N1);
    gs.N2 = P(sd.N2);
    const P N3 = P(sd.N3), N4 = P(sd.N4), N7 = P(sd.N7);
    V3<P> ut = {P(0), N3 * cb.tb1.z + N4 * cb.tb2.z, (-N3) * cb.tb1.y + (-N4) * cb.tb2.y};
    gs.Sut = skew(ut);
    gs.N7l = N7 / cb.ln;
    gs.N7l2 = gs.N7l / cb.ln;
    gs.PP1 = rows(P(0), P(0), P(0), P(0), gs.N7l, P(0), P(0), P(0), gs.N7l);
    gs.PP2 = rows(P(0), P(0), P(0), P(0), P(0), N3, P(0), -N3, P(0));
    gs.PP4 = rows(P(0), P(0), P(0), P(0), P(0), N4, P(0), -N4, P(0));
    return gs;
}
// H₁ᵃ (compute&assemble_interactions_beams.jl:84-87), with P₁P³ = -P₁P¹ and Gᵀ³ = -Gᵀ¹
template <typename P, int A> EBD M3<P> gauss_H1(const ContactBeam<P>& cb, const GaussShape<P>& gs) {
    const M3<P> I = eye<P>();
    if constexpr (A == 0) return (gs.N1 * I + gs.PP1) - gs.Sut * cb.G1;
    else if constexpr (A == 1) return gs.PP2 - gs.Sut * cb.G2;
    else if constexpr (A == 2) return (gs.N2 * I + (-gs.PP1)) - gs.Sut * (-cb.G1);
    else return gs.PP4 - gs.Sut * cb.G4;
}

// ηₙ after the smoothstep of a contacting point with gap gₙ (unchanged where gₙ ≤ 0)
template <typename P> EBD P eta_after(P eta_n, P gn, P gb) {
    if (gn <= P(0)) return eta_n;
    P ss = (gn - gb) / gb;
    return ((ss * ss) * (P(3) + P(2) * ss)) * eta_n;
}

// Phase 1: Gauss point ig of one beam, up to 𝒦ᶠᶜᵇ, 𝒞ᶠᶜᵇ, 𝓕ᶜ (and, with FORCE, the point's terms
// of Tᶜ and of the contact energy) in LDS slot gi. eta_n: ηₙ after the beam's earlier contacting
// points; gn, n, Hg, h1, H1F1: the point's gap and kinematics, computed before the barrier that
// published the gaps.
template <typename P, typename C, bool FORCE>
EBD void contact_point(const ContactBeam<P>& cb, const NodeState& ns, const GaussPts& gp, const ContactArgs<C>& ca, int ig,
                       int gi, int lane, P eta_n, P gn, const V3<P>& n, const M3<P>& Hg, const V3<P>& h1,
                       const M3<P>& H1F1, ContactShared<P, FORCE>& sh) {
    const GaussShape<P> gs = gauss_shape(cb, shape_fn(cb.l0, gp.z[ig]));
    M3<P> H1[4];
    H1[0] = gauss_H1<P, 0>(cb, gs);
    H1[1] = gauss_H1<P, 1>(cb, gs);
    H1[2] = gauss_H1<P, 2>(cb, gs);
    H1[3] = gauss_H1<P, 3>(cb, gs);
    V3<P> Ud[4];
    Ud[0] = tmul(cb.Re, cvt<P>(ldv(ns.ud, cb.na)));
    Ud[1] = tmul(cb.Re, cvt<P>(ldv(ns.wd, cb.na)));
    Ud[2] = tmul(cb.Re, cvt<P>(ldv(ns.ud, cb.nb)));
    Ud[3] = tmul(cb.Re, cvt<P>(ldv(ns.wd, cb.nb)));
    const M3<P> I = eye<P>();
    const M3<P>& Re = cb.Re;
    const M3<P>& ReT = cb.ReT;
    const M3<P> G3 = -cb.G1;
    const M3<P> G1R = cb.G1 * ReT, G2R = cb.G2 * ReT, G4R = cb.G4 * ReT;
    const P w = P(gp.w[ig]);
    const P rb = P(cb.rb);
    const P kn = P(ca.kn), mu = P(ca.mu), eps_t = P(ca.eps_t);
    const V3<P> ud0 = Re * h1;
    // regularize_gap_penalty_beams
    P gb = rb / P(4), hh = gb / P(2), pn, dpn, Pi;
    if (gn <= P(0)) {
        pn = hh - gn; dpn = P(-1);
        Pi = ((gn * gn) / P(2) - hh * gn) + (gb * gb) / P(6);
    } else {
        P aux = (gb - hh) / (gb * gb);
        pn = (aux * (gn * gn) - gn) + hh;
        dpn = (P(2) * aux) * gn - P(1);
        Pi = (((gb - hh) / (P(3) * (gb * gb))) * ((gn * gn) * gn) - (gn * gn) / P(2) + hh * gn) - (gb * gb) / P(6);
    }
    // smoothstep_transition_beams
    P deta;
    if (gn <= P(0)) {
        deta = P(0);
    } else {
        P ss = (gn - gb) / gb;
        deta = (((P(6) / gb) * ss) * (P(1) + ss)) * eta_n;
        eta_n = ((ss * ss) * (P(3) + P(2) * ss)) * eta_n;
    }
    P udn_mag = dot(ud0, n);
    V3<P> udn = udn_mag * n;
    V3<P> udt = ud0 - udn;
    P udt2 = dot(udt, udt);
    if constexpr (FORCE) sh.en[gi][lane] = double((w * kn) * Pi);
    V3<P> fn = (kn * pn) * n - eta_n * udn;
    P mureg = mu / sqrt(udt2 + eps_t);
    V3<P> ft = (((-kn) * pn) * mureg) * udt;
    V3<P> fc = fn + ft;
    V3<P> Fc = tmul(Re, fc);
    sh.fc[gi][0][lane] = Fc.x;
    sh.fc[gi][1][lane] = Fc.y;
    sh.fc[gi][2][lane] = Fc.z;
    M3<P> RHR[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) RHR[a] = (Re * H1[a]) * ReT;
    if constexpr (FORCE) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            V3<P> t = w * tmul(RHR[a], fc);
            sh.tc[gi][3 * a][lane] = double(t.x);
            sh.tc[gi][3 * a + 1][lane] = double(t.y);
            sh.tc[gi][3 * a + 2][lane] = double(t.z);
        }
    }
    M3<P> aux = dot(n, ud0) * Hg + outer(n, Hg * ud0);
    M3<P> RSh1 = Re * skew(h1);
    const V3<P>& Ud1 = Ud[0];
    const V3<P>& Wd1 = Ud[1];
    const V3<P>& Ud2 = Ud[2];
    const V3<P>& Wd2 = Ud[3];
    const V3<P>& v1 = cb.v1;
    M3<P> A1Dr = rows(P(0), P(0), P(0), v1.x * (Ud1.y - Ud2.y), v1.y * (Ud1.y - Ud2.y), v1.z * (Ud1.y - Ud2.y),
                      v1.x * (Ud1.z - Ud2.z), v1.y * (Ud1.z - Ud2.z), v1.z * (Ud1.z - Ud2.z));
    M3<P> RSGD = Re * skew(((cb.G1 * Ud1 + cb.G2 * Wd1) + G3 * Ud2) + cb.G4 * Wd2);
    M3<P> RH1SD = Re * H1F1;
    M3<P> du0dd[4];
    du0dd[0] = ((-(RSh1 * G1R)) + ((gs.N7l2 * Re) * A1Dr + (RSGD * gs.PP1) * ReT)) + RH1SD * G1R;
    du0dd[1] = ((-(RSh1 * G2R)) + (RSGD * gs.PP2) * ReT) + RH1SD * G2R;
    du0dd[2] = -du0dd[0];
    du0dd[3] = ((-(RSh1 * G4R)) + (RSGD * gs.PP4) * ReT) + RH1SD * G4R;
    M3<P> nn = outer(n, n);
    M3<P> auxn = ((kn * dpn) * nn + (kn * pn) * Hg) - outer(deta * udn, n);
    M3<P> aux1 = outer((((-mureg) * kn) * dpn) * udt, n);
    M3<P> aux2 = (((-mureg) * kn) * pn) * (I - outer((P(1) / (udt2 + eps_t)) * udt, udt));
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        M3<P> dundd = aux * RHR[b] + nn * du0dd[b];
        M3<P> dutdd = du0dd[b] - dundd;
        M3<P> dutdv = RHR[b] - nn * RHR[b];
        M3<P> Kfn = auxn * RHR[b] - eta_n * dundd;
        M3<P> Cfn = ((-eta_n) * nn) * RHR[b];
        M3<P> Kft = aux1 * RHR[b] + aux2 * dutdd;
        M3<P> Cft = aux2 * dutdv;
        const M3<P> Kfc = Kfn + Kft;
        const M3<P> Cfc = Cfn + Cft;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            sh.kfc[gi][9 * b + k][lane] = Kfc.a[k];
            sh.cfc[gi][9 * b + k][lane] = Cfc.a[k];
        }
    }
}

// Phase 2, block row A of one beam: Kᶜᴬᵇ += ωᴳ (t₁ + t₂ + t₃ + t₄), Cᶜᴬᵇ += ωᴳ RₑH₁ᴬRₑᵀᵀ 𝒞ᶠᶜᵇ over the
// chunk's contacting points (compute&assemble_interactions_beams.jl:132-212, :308-372)
template <typename P, bool FORCE, int A>
EBD void contact_row(const ContactBeam<P>& cb, const GaussPts& gp, int g0, int lane, const ContactShared<P, FORCE>& sh,
                     M3<P> (&Kc)[4], M3<P> (&Cc)[4], V3<P>& Tca, double& energy) {
    const M3<P>& Re = cb.Re;
    const M3<P>& ReT = cb.ReT;
    const M3<P> G1R = cb.G1 * ReT, G2R = cb.G2 * ReT, G4R = cb.G4 * ReT;
    const M3<P> RG1 = Re * tr(cb.G1);
    const M3<P> RGa = A == 0 ? RG1 : (A == 1 ? Re * tr(cb.G2) : (A == 2 ? -RG1 : Re * tr(cb.G4)));
    const V3<P>& v1 = cb.v1;
    const int gend = min(gp.n, g0 + CONTACT_GP_CHUNK);
#pragma unroll 1
    for (int ig = g0; ig < gend; ++ig) {
        if (!(cb.mask & (1u << ig))) continue;
        const int gi = ig - g0;
        const P w = P(gp.w[ig]);
        const GaussShape<P> gs = gauss_shape(cb, shape_fn(cb.l0, gp.z[ig]));
        const M3<P> H1 = gauss_H1<P, A>(cb, gs);
        const V3<P> Fc = {sh.fc[gi][0][lane], sh.fc[gi][1][lane], sh.fc[gi][2][lane]};
        if constexpr (FORCE) {
            Tca += V3<P>{P(sh.tc[gi][3 * A][lane]), P(sh.tc[gi][3 * A + 1][lane]), P(sh.tc[gi][3 * A + 2][lane])};
            if (A == 0) energy += sh.en[gi][lane];
        }
        const M3<P> tRHR = tr((Re * H1) * ReT);
        const M3<P> SF = skew(Fc);
        const M3<P> RS = Re * skew(tmul(H1, Fc));
        const M3<P> RHS = (Re * tr(H1)) * SF;
        const M3<P> SFP1 = (SF * gs.PP1) * ReT, SFP2 = (SF * gs.PP2) * ReT, SFP4 = (SF * gs.PP4) * ReT;
        const M3<P> A1F = rows(P(0), P(0), P(0), Fc.y * v1.x, Fc.y * v1.y, Fc.y * v1.z, Fc.z * v1.x, Fc.z * v1.y, Fc.z * v1.z);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            M3<P> Kfc, Cfc;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                Kfc.a[k] = sh.kfc[gi][9 * b + k][lane];
                Cfc.a[k] = sh.cfc[gi][9 * b + k][lane];
            }
            // t₁ᵃᵇ = -Rₑ S(H₁ᵃᵀ𝓕ᶜ) Gᵇ Rₑᵀ with G³ = -G¹; t₃ᵃᵇ = Rₑ H₁ᵃᵀ S𝓕ᶜ Gᵇ Rₑᵀ
            const M3<P> GbR = b == 0 ? G1R : (b == 1 ? G2R : (b == 2 ? G1R : G4R));
            const M3<P> t1 = b == 2 ? RS * GbR : -(RS * GbR);
            const M3<P> t3 = b == 2 ? -(RHS * GbR) : RHS * GbR;
            // t₂ (:166-184): rows 1 and 3 carry the A₁ᵀ𝓕ᶜr term; row 3 mirrors row 1
            const M3<P> SFPb = b == 0 ? SFP1 : (b == 1 ? SFP2 : (b == 2 ? SFP1 : SFP4));
            M3<P> t2;
            if constexpr (A == 0 || A == 2) {
                M3<P> base;
                if (b == 0 || b == 2) base = (gs.N7l2 * Re) * A1F - RG1 * SFP1;
                else base = -(RG1 * SFPb);
                const bool neg = (b == 2) != (A == 2);
                t2 = neg ? -base : base;
            } else {
                const M3<P> base = -(RGa * SFP1);
                if (b == 0) t2 = base;
                else if (b == 2) t2 = -base;
                else t2 = -(RGa * SFPb);
            }
            const M3<P> t4 = tRHR * Kfc;
            Kc[b] += w * (((t1 + t2) + t3) + t4);
            Cc[b] += (w * tRHR) * Cfc;
        }
    }
}

// the end of block row A (the wavefront's row, runtime but uniform): l₀/2 and, on the C rows of the
// rotational dofs, Tₛ⁻¹(Θ)ᵀ (:411-430); Kc, Cc become the values the staging receives (K -= Kᶜ,
// C -= (1+α)Cᶜ, assemble_contact!). With FORCE also Tᶜ of slot A and (row 0) the contact energy.
template <typename P, bool FORCE>
EBD void contact_row_finish(const ContactBeam<P>& cb, const ContactArgs<P>& ca, int A, M3<P> (&Kc)[4],
                            M3<P> (&Cc)[4], const V3<P>& Tca, double energy, const M3<P>& TTt) {
    const int64_t e = cb.e;
    const P l02 = P(cb.l0 / 2);
    if constexpr (FORCE) {
        if (A == 0) ca.energy[3 * e + 2] = double(l02) * energy;
        V3<P> t = l02 * Tca;
        ca.Tc[e * 12 + 3 * A] = double(t.x);
        ca.Tc[e * 12 + 3 * A + 1] = double(t.y);
        ca.Tc[e * 12 + 3 * A + 2] = double(t.z);
    }
    const P c1a = P(1 + ca.alpha);
    const bool rot = A == 1 || A == 3;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        M3<P> Cm = l02 * Cc[b];
        if (rot) Cm = Cm * TTt;
        Kc[b] = -(l02 * Kc[b]);
        Cc[b] = -(c1a * Cm);
    }
}

// A tile's 12x12 contributions pass through LDS, one record of 144 entries per beam in the staging's
// order (stride 148: 16-byte aligned rows), so that the read-modify-write of the staging moves whole
// 16-byte pieces of one contiguous record per beam instead of 3x3 blocks scattered over 64 beams.
constexpr int CONTACT_REC = 148;
template <typename P> EBD void contact_row_to_lds(P* buf, int lane, int A, const M3<P> (&V)[4]) {
    P* r = buf + lane * CONTACT_REC;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int R = 3 * A + i, Cl = 3 * b + j;
                r[(2 * (R / 6) + Cl / 6) * 36 + (R % 6) * 6 + Cl % 6] = V[b].a[i + 3 * j];
            }
}
// 16-byte pieces: 36 per FP32 record, 72 per FP64 record; thread t takes pieces t, t + 256, ...
template <typename P> struct Piece { P v[16 / sizeof(P)]; };
template <typename P> constexpr int pieces_per_rec() { return 144 * (int)sizeof(P) / 16; }
template <typename P> constexpr int pieces_per_thread() { return CONTACT_TILE * pieces_per_rec<P>() / (4 * CONTACT_TILE); }
template <typename P>
EBD void contact_tile_load(const P* __restrict__ src, const int64_t* tile_e, int nvalid, Piece<P> (&v)[pieces_per_thread<P>()]) {
    constexpr int PR = pieces_per_rec<P>(), NP = 16 / sizeof(P);
    const int total = nvalid * PR;
#pragma unroll
    for (int u = 0; u < pieces_per_thread<P>(); ++u) {
        const int q = threadIdx.x + u * 4 * CONTACT_TILE;
        const int j = q / PR, k = q - PR * j;
        if (q < total) v[u] = *reinterpret_cast<const Piece<P>*>(src + tile_e[j] * 144 + NP * k);
    }
}
template <typename P>
EBD void contact_tile_store(P* __restrict__ dst, const P* buf, const int64_t* tile_e, int nvalid,
                            const Piece<P> (&v)[pieces_per_thread<P>()]) {
    constexpr int PR = pieces_per_rec<P>(), NP = 16 / sizeof(P);
    const int total = nvalid * PR;
#pragma unroll
    for (int u = 0; u < pieces_per_thread<P>(); ++u) {
        const int q = threadIdx.x + u * 4 * CONTACT_TILE;
        const int j = q / PR, k = q - PR * j;
        if (q < total) {
            Piece<P> w;
#pragma unroll
            for (int c = 0; c < NP; ++c) w.v[c] = v[u].v[c] + buf[j * CONTACT_REC + NP * k + c];
            *reinterpret_cast<Piece<P>*>(dst + tile_e[j] * 144 + NP * k) = w;
        }
    }
}

// The tile's LDS: the points' records, then the tile's contributions to the staging (FP32: K and C
// at once; FP64: K, then C, in one record buffer, so that the FP64 tile stays within 146 KB)
template <typename P> struct ContactLDS {
    static constexpr bool FORCE = sizeof(P) == sizeof(double);
    static constexpr int NREC = FORCE ? 1 : 2;
    union {
        ContactShared<P, FORCE> pts;
        P rec[NREC][CONTACT_TILE * CONTACT_REC];
    } u;
    int64_t tile_e[CONTACT_TILE];
};

// One 64-beam tile of the contact list (at most CONTACT_GP_CHUNK Gauss points per beam: the
// reference's rule has three, ωᴳ_beams is a Vec3): phase 1 (wavefront g: point g; the wavefront left
// over: the Tₛ⁻¹ blocks of the C rows 2 and 4), phase 2 (wavefront a: block row a), the sums folded
// into the staging through LDS. FP64 form (P = double): the contact force and energy too; precision
// form (P = float): the tangent alone (contact_force_kernel computes the FP64 force). Every thread
// of a 256-thread workgroup calls it for the same tile.
template <typename P>
EBD void contact_tile(int64_t tile, int32_t nlisted, const NodeState& ns, const BeamConsts& bc, const BeamStage<P>& st,
                      const GaussPts& gp, const ContactArgs<P>& ca, const uint64_t* __restrict__ list,
                      ContactLDS<P>& lds) {
    constexpr bool FORCE = sizeof(P) == sizeof(double);
    auto& shu = lds.u;
    int64_t* tile_e = lds.tile_e;
    ContactShared<P, FORCE>& sh = shu.pts;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / CONTACT_TILE);
    const int lane = threadIdx.x % CONTACT_TILE;
    const int64_t li = tile * CONTACT_TILE + lane;
    if (tile * CONTACT_TILE >= nlisted) return;   // uniform over the workgroup
    if (wv == 0 && ca.stats) {
        // the work of this launch: listed beams and Gauss points in contact (one wavefront's sums)
        int nb = li < nlisted ? 1 : 0;
        int ng = li < nlisted ? __popc((unsigned)(list[li] & 0xffu)) : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            nb += __shfl_down(nb, off, 64);
            ng += __shfl_down(ng, off, 64);
        }
        if (threadIdx.x == 0 && nb > 0) {
            atomicAdd(&ca.stats[0], (unsigned long long)nb);
            atomicAdd(&ca.stats[1], (unsigned long long)ng);
        }
    }
    const bool valid = li < nlisted;
    const uint64_t ent = valid ? list[li] : 0;
    const ContactBeam<P> cb = contact_beam<P>((int64_t)(ent >> 8), (unsigned)(ent & 0xffu), ns, bc, ca);
    const int nvalid = (int)min<int64_t>(CONTACT_TILE, nlisted - tile * CONTACT_TILE);
    if (wv == 0) tile_e[lane] = cb.e;
    const P gb = P(cb.rb) / P(4);
    {
        const int g0 = 0, ig = wv;
        const bool mine = wv < CONTACT_GP_CHUNK && valid && ig < gp.n && (cb.mask & (1u << ig));
        // phase 1a: the point's kinematics and gap
        P gn = P(0);
        V3<P> n, h1;
        M3<P> Hg, H1F1;
        if (mine) {
            const GaussShape<P> gs = gauss_shape(cb, shape_fn(cb.l0, gp.z[ig]));
            const ShapeF sd = shape_fn(cb.l0, gp.z[ig]);
            V3<double> utd = {0.0, sd.N3 * cb.Tb1.z + sd.N4 * cb.Tb2.z, (-sd.N3) * cb.Tb1.y + (-sd.N4) * cb.Tb2.y};
            V3<double> xG = (sd.N1 * cb.x1 + sd.N2 * cb.x2) + cb.Red * utd;
            const M3<P> H10 = gauss_H1<P, 0>(cb, gs), H11 = gauss_H1<P, 1>(cb, gs);
            const M3<P> H12 = gauss_H1<P, 2>(cb, gs), H13 = gauss_H1<P, 3>(cb, gs);
            const V3<P> Ud1 = tmul(cb.Re, cvt<P>(ldv(ns.ud, cb.na))), Wd1 = tmul(cb.Re, cvt<P>(ldv(ns.wd, cb.na)));
            const V3<P> Ud2 = tmul(cb.Re, cvt<P>(ldv(ns.ud, cb.nb))), Wd2 = tmul(cb.Re, cvt<P>(ldv(ns.wd, cb.nb)));
            h1 = ((H10 * Ud1 + H11 * Wd1) + H12 * Ud2) + H13 * Wd2;
            H1F1 = ((H10 * skew(Ud1) + H11 * skew(Wd1)) + H12 * skew(Ud2)) + H13 * skew(Wd2);
            contact_gap<P, P>(ca, cb.inst, xG, cb.rb, gn, n, Hg);
            sh.gn[wv][lane] = gn;
        } else if (wv == CONTACT_GP_CHUNK && valid) {
            const M3<P> T1 = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, cb.na)))));
            const M3<P> T2 = cvt<P>(tr(Ts_inv(toangle(ldm(ns.dR, cb.nb)))));
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                sh.tt[0][k][lane] = T1.a[k];
                sh.tt[1][k][lane] = T2.a[k];
            }
        }
        __syncthreads();
        // phase 1b: ηₙ after the earlier contacting points, then the point's force derivatives
        if (mine) {
            P eta_n = P(ca.eta_n);
            for (int j = g0; j < ig; ++j)
                if (cb.mask & (1u << j)) eta_n = eta_after(eta_n, sh.gn[j - g0][lane], gb);
            contact_point<P, P, FORCE>(cb, ns, gp, ca, ig, wv, lane, eta_n, gn, n, Hg, h1, H1F1, sh);
        }
        __syncthreads();
        // phase 2: block row wv over the chunk's points, folded into the staging
        M3<P> Kc[4], Cc[4];
        if (valid) {
            for (int b = 0; b < 4; ++b) { Kc[b] = zm<P>(); Cc[b] = zm<P>(); }
            V3<P> Tca = zv<P>();
            double energy = 0;
            if (wv == 0) contact_row<P, FORCE, 0>(cb, gp, g0, lane, sh, Kc, Cc, Tca, energy);
            else if (wv == 1) contact_row<P, FORCE, 1>(cb, gp, g0, lane, sh, Kc, Cc, Tca, energy);
            else if (wv == 2) contact_row<P, FORCE, 2>(cb, gp, g0, lane, sh, Kc, Cc, Tca, energy);
            else contact_row<P, FORCE, 3>(cb, gp, g0, lane, sh, Kc, Cc, Tca, energy);
            M3<P> TTt = eye<P>();
            if (wv == 1 || wv == 3)
                for (int k = 0; k < 9; ++k) TTt.a[k] = sh.tt[wv == 1 ? 0 : 1][k][lane];
            contact_row_finish<P, FORCE>(cb, ca, wv, Kc, Cc, Tca, energy, TTt);
        }
        // the staging records the tile adds to, all in flight at once
        Piece<P> oldK[pieces_per_thread<P>()], oldC[pieces_per_thread<P>()];
        contact_tile_load(st.K, tile_e, nvalid, oldK);
        contact_tile_load(st.C, tile_e, nvalid, oldC);
        __syncthreads();   // the points' records are read: LDS takes the tile's K and C contributions
        constexpr int KR = 0, CR = ContactLDS<P>::NREC - 1;
        if (valid) {
            contact_row_to_lds(shu.rec[KR], lane, wv, Kc);
            if (CR != KR) contact_row_to_lds(shu.rec[CR], lane, wv, Cc);
        }
        __syncthreads();
        contact_tile_store(st.K, shu.rec[KR], tile_e, nvalid, oldK);
        if (CR == KR) {
            __syncthreads();
            if (valid) contact_row_to_lds(shu.rec[CR], lane, wv, Cc);
            __syncthreads();
        }
        contact_tile_store(st.C, shu.rec[CR], tile_e, nvalid, oldC);
        __syncthreads();
    }
}

template <typename P>
__global__ void __launch_bounds__(4 * CONTACT_TILE) contact_kernel(NodeState ns, BeamConsts bc, BeamStage<P> st,
                                                                   GaussPts gp, ContactArgs<P> ca,
                                                                   const uint64_t* __restrict__ list,
                                                                   const int32_t* __restrict__ count) {
    __shared__ ContactLDS<P> lds;
    contact_tile<P>(blockIdx.x, *count, ns, bc, st, gp, ca, list, lds);
}

}  // namespace ebd
