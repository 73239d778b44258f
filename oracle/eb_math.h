// eb_math.h — TEST INFRASTRUCTURE (oracle). Never linked into the product library.
//
// 3-vector / 3x3-matrix algebra with the evaluation order of StaticArrays as the reference uses
// it: sums run left to right over the inner index, n-ary `*` on matrices folds left
// ((A*B)*C), matrix*matrix*vector is A*(B*x) (LinearAlgebra's 3-argument method), Diagonal
// operands behave like dense ones with exact zeros. Built with -ffp-contract=off so that no
// multiply-add is fused. Column-major storage like SMatrix{3,3}.
#pragma once
#include <cmath>

#ifndef EBO_REAL
#define EBO_REAL double
#endif

namespace ebo {

// scalar type of the restatement: double, or an operation-counting wrapper (flopcount.cpp)
typedef EBO_REAL real;
using std::acos;
using std::cos;
using std::fabs;
using std::floor;
using std::sin;
using std::sqrt;

static const real EPS = 2.220446049250313e-16;  // eps(Float64)

struct V3 {
    real v[3];
    real& operator[](int i) { return v[i]; }
    real operator[](int i) const { return v[i]; }
};
struct M3 {
    real m[9];  // m[i + 3*j] = A[i,j]
    real& operator()(int i, int j) { return m[i + 3 * j]; }
    real operator()(int i, int j) const { return m[i + 3 * j]; }
};

inline V3 v3(real a, real b, real c) { return V3{{a, b, c}}; }
inline V3 zv() { return v3(0, 0, 0); }
inline M3 zm() { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = 0.0; return r; }
inline M3 eye() { M3 r = zm(); r(0, 0) = 1; r(1, 1) = 1; r(2, 2) = 1; return r; }
// row-major literal, like @SMatrix [a b c; d e f; g h i]
inline M3 rows(real a, real b, real c, real d, real e, real f, real g, real h, real i) {
    M3 r; r(0, 0) = a; r(0, 1) = b; r(0, 2) = c; r(1, 0) = d; r(1, 1) = e; r(1, 2) = f;
    r(2, 0) = g; r(2, 1) = h; r(2, 2) = i; return r;
}
inline M3 diag(real a, real b, real c) { M3 r = zm(); r(0, 0) = a; r(1, 1) = b; r(2, 2) = c; return r; }

inline V3 operator+(const V3& a, const V3& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 operator-(const V3& a, const V3& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 operator-(const V3& a) { return v3(-a[0], -a[1], -a[2]); }
inline V3 operator*(real s, const V3& a) { return v3(s * a[0], s * a[1], s * a[2]); }
inline V3 operator*(const V3& a, real s) { return v3(a[0] * s, a[1] * s, a[2] * s); }
inline V3 operator/(const V3& a, real s) { return v3(a[0] / s, a[1] / s, a[2] / s); }
inline V3& operator+=(V3& a, const V3& b) { a = a + b; return a; }

inline M3 operator+(const M3& a, const M3& b) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] + b.m[k]; return r; }
inline M3 operator-(const M3& a, const M3& b) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] - b.m[k]; return r; }
inline M3 operator-(const M3& a) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = -a.m[k]; return r; }
inline M3 operator*(real s, const M3& a) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = s * a.m[k]; return r; }
inline M3 operator*(const M3& a, real s) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] * s; return r; }
inline M3 operator/(const M3& a, real s) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] / s; return r; }
inline M3& operator+=(M3& a, const M3& b) { a = a + b; return a; }

inline M3 operator*(const M3& a, const M3& b) {
    M3 r;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
    return r;
}
inline V3 operator*(const M3& a, const V3& x) {
    V3 r;
    for (int i = 0; i < 3; ++i) r[i] = a(i, 0) * x[0] + a(i, 1) * x[1] + a(i, 2) * x[2];
    return r;
}
inline M3 tr(const M3& a) { M3 r; for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) r(i, j) = a(j, i); return r; }
inline real trace(const M3& a) { return a(0, 0) + a(1, 1) + a(2, 2); }
inline real dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline real norm(const V3& a) { return sqrt(dot(a, a)); }
inline V3 cross(const V3& a, const V3& b) {
    return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline M3 outer(const V3& a, const V3& b) { M3 r; for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) r(i, j) = a[i] * b[j]; return r; }
inline M3 skew(const V3& v) { return rows(0, -v[2], v[1], v[2], 0, -v[0], -v[1], v[0], 0); }
inline V3 col(const M3& a, int j) { return v3(a(0, j), a(1, j), a(2, j)); }
inline M3 from_cols(const V3& a, const V3& b, const V3& c) {
    M3 r; for (int i = 0; i < 3; ++i) { r(i, 0) = a[i]; r(i, 1) = b[i]; r(i, 2) = c[i]; } return r;
}

// ---- rotation utilities: src/beams/utils_beams.jl ------------------------------------------

// rotation_matrix, utils_beams.jl:11-32 (Rodrigues; identity when ‖Θ‖ <= 10ε)
inline M3 rotation_matrix(const V3& th) {
    real n = norm(th);
    if (n > 10 * EPS) {
        real s = sin(n), c = cos(n);
        M3 S = skew(th);
        return (eye() + (s / n) * S) + ((1 - c) / (n * n) * S) * S;
    }
    return eye();
}

// toangle, utils_beams.jl:90-104 (log map through acos of the trace)
inline V3 toangle(const M3& R) {
    real t = trace(R);
    real nv;
    if (fabs((t - 1) / 2) < 1) {
        nv = acos((t - 1) / 2);
        if (nv < 10 * EPS) nv = 10 * EPS;
    } else {
        nv = 10 * EPS;
    }
    real c = 1 / (2 * sin(nv));
    V3 a = v3(R(2, 1) - R(1, 2), R(0, 2) - R(2, 0), R(1, 0) - R(0, 1));
    return nv * (c * a);
}

// Tₛ⁻¹, utils_beams.jl:107-122
inline M3 Ts_inv(const V3& th) {
    real n = norm(th);
    if (n < 10 * EPS) return eye();
    M3 S = skew(th);
    real s2 = sin(n / 2), c2 = cos(n / 2);
    real coef = (s2 - n / 2 * c2) / (s2 * (n * n));
    return (eye() - S / 2) + (coef * S) * S;
}

// Tₛ, utils_beams.jl:125-140 (only used by the oracle's own checks: Tₛ⁻¹·Tₛ = I)
inline M3 Ts(const V3& th) {
    real n = norm(th);
    if (n < 10 * EPS) return eye();
    M3 S = skew(th);
    real a = sin(n / 2) / n;
    return (eye() + (2 * (a * a)) * S) + ((1 - sin(n) / n) / (n * n)) * (S * S);
}

// compute_η_μ, utils_beams.jl:180-198
inline void eta_mu(const V3& tb, real& eta, real& mu) {
    real t = norm(tb);
    if (t < 10 * EPS) { eta = 1.0 / 12; mu = 1.0 / 360; return; }
    real s = sin(t), c = cos(t), s2 = sin(t / 2);
    eta = (2 * s - t * (1 + c)) / ((2 * (t * t)) * s);
    real t4 = (t * t) * (t * t);
    mu = (t * (t + s) - 8 * (s2 * s2)) / ((4 * t4) * (s2 * s2));
}

// compute_K̄ₕ, utils_beams.jl:201-215
inline M3 Kh_bar(const V3& tb, const V3& Mb, const M3& Tinv, real eta, real mu) {
    M3 TM = outer(tb, Mb);
    M3 MT = tr(TM);
    M3 St = skew(tb), SM = skew(Mb);
    M3 a = (TM - 2 * MT) + dot(tb, Mb) * eye();
    M3 b = (St * St) * MT;
    return ((eta * a + mu * b) - SM / 2) * Tinv;
}

// local_Rₑ_and_aux, utils_beams.jl:35-87
struct LocalFrame {
    M3 Re;
    V3 r1, r3;   // r¹ = -v₁ᵀ, r³ = v₁ᵀ (stored as vectors)
    real eta;
    M3 G1, G2, G3, G4;  // Gᵀu₁, GᵀΘ₁, Gᵀu₂, GᵀΘ₂
    M3 D3;
};
inline LocalFrame local_Re_and_aux(const V3& x1, const V3& x2, const M3& R1, const M3& R2,
                                   const V3& Re0E2, real ln) {
    LocalFrame f;
    V3 v1 = (x2 - x1) / ln;
    V3 p1 = R1 * Re0E2, p2 = R2 * Re0E2;
    V3 p = (p1 + p2) / 2;
    V3 v3_ = cross(v1, p);
    v3_ = v3_ / norm(v3_);
    V3 v2 = cross(v3_, v1);
    f.Re = from_cols(v1, v2, v3_);
    f.r1 = -v1;
    f.r3 = v1;
    real q1 = dot(v1, p), q2 = dot(v2, p);
    real q11 = dot(v1, p1), q12 = dot(v2, p1);
    real q21 = dot(v1, p2), q22 = dot(v2, p2);
    real eta = q1 / q2, eta11 = q11 / q2, eta12 = q12 / q2, eta21 = q21 / q2, eta22 = q22 / q2;
    f.eta = eta;
    f.G1 = rows(0, 0, eta / ln, 0, 0, 1 / ln, 0, -1 / ln, 0);
    f.G3 = -f.G1;
    f.G2 = rows(eta12 / 2, -eta11 / 2, 0, 0, 0, 0, 0, 0, 0);
    f.G4 = rows(eta22 / 2, -eta21 / 2, 0, 0, 0, 0, 0, 0, 0);
    f.D3 = (eye() - outer(v1, v1)) / ln;
    return f;
}

// get_Rₑ⁰, src/beams/def_beam.jl:93-116
inline M3 get_Re0(const V3& X1, const V3& X2) {
    real l0 = norm(X2 - X1);
    V3 E1 = v3(1, 0, 0);
    V3 E10 = (X2 - X1) / l0;
    V3 v = cross(E1, E10);
    real s = norm(v), c = dot(E1, E10);
    if (c < -(1 - 2 * EPS)) return rows(-1, 0, 0, 0, 1, 0, 0, 0, -1);
    if (c > (1 - 2 * EPS)) return eye();
    M3 Sv = skew(v);
    return (eye() + Sv) + ((1 - c) / (s * s) * Sv) * Sv;
}

}  // namespace ebo
