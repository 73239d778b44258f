// eb_dmath.h — device 3-vector / 3x3 algebra and corotational rotation utilities for gfx950.
//
// Evaluation order follows the reference's StaticArrays expressions (left-to-right inner sums,
// left-folded n-ary products, A*(B*x) for matrix*matrix*vector); with contraction off this makes
// the rotation extraction bit-reproducible, which matters because acos((tr R - 1)/2) amplifies
// input rounding by 1/θ² at small angles (src/beams/utils_beams.jl:90-104).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ebd {

template <typename T> struct V3 { T x, y, z; };
template <typename T> struct M3 { T a[9]; };  // column-major: a[i + 3 j]

#define EBD __device__ __forceinline__

template <typename T> EBD T& at(M3<T>& m, int i, int j) { return m.a[i + 3 * j]; }
template <typename T> EBD T at(const M3<T>& m, int i, int j) { return m.a[i + 3 * j]; }

template <typename T> EBD V3<T> v3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <typename T> EBD V3<T> zv() { return V3<T>{T(0), T(0), T(0)}; }
template <typename T> EBD M3<T> zm() { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = T(0); return r; }
template <typename T> EBD M3<T> eye() { M3<T> r = zm<T>(); r.a[0] = T(1); r.a[4] = T(1); r.a[8] = T(1); return r; }
template <typename T> EBD M3<T> rows(T a, T b, T c, T d, T e, T f, T g, T h, T i) {
    M3<T> r; r.a[0] = a; r.a[3] = b; r.a[6] = c; r.a[1] = d; r.a[4] = e; r.a[7] = f; r.a[2] = g; r.a[5] = h; r.a[8] = i;
    return r;
}
template <typename T> EBD M3<T> diag(T a, T b, T c) { M3<T> r = zm<T>(); r.a[0] = a; r.a[4] = b; r.a[8] = c; return r; }

template <typename T> EBD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> EBD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> EBD V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> EBD V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> EBD V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> EBD V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> EBD V3<T>& operator+=(V3<T>& a, V3<T> b) { a = a + b; return a; }

template <typename T> EBD M3<T> operator+(const M3<T>& a, const M3<T>& b) { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = a.a[k] + b.a[k]; return r; }
template <typename T> EBD M3<T> operator-(const M3<T>& a, const M3<T>& b) { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = a.a[k] - b.a[k]; return r; }
template <typename T> EBD M3<T> operator-(const M3<T>& a) { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = -a.a[k]; return r; }
template <typename T> EBD M3<T> operator*(T s, const M3<T>& a) { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = s * a.a[k]; return r; }
template <typename T> EBD M3<T> operator*(const M3<T>& a, T s) { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = a.a[k] * s; return r; }
template <typename T> EBD M3<T> operator/(const M3<T>& a, T s) { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = a.a[k] / s; return r; }
template <typename T> EBD M3<T>& operator+=(M3<T>& a, const M3<T>& b) { a = a + b; return a; }

// a·b + c·d + e·f, the inner sum of every 3x3 product. FP64: two roundings per product term in the
// reference's left-to-right order (contraction is off in this library). FP32, the precision form's
// matrices: the compiler may fuse each multiply with the following add, one rounding per term
// and a third fewer instructions; FP32 results need not reproduce the reference's bits.
template <typename T> EBD T sum3(T a, T b, T c, T d, T e, T f) {
    if constexpr (sizeof(T) == sizeof(float)) {
#pragma clang fp contract(fast)
        return a * b + c * d + e * f;
    } else {
        return a * b + c * d + e * f;
    }
}
template <typename T> EBD M3<T> operator*(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            r.a[i + 3 * j] = sum3(a.a[i], b.a[3 * j], a.a[i + 3], b.a[3 * j + 1], a.a[i + 6], b.a[3 * j + 2]);
    return r;
}
template <typename T> EBD V3<T> operator*(const M3<T>& a, V3<T> x) {
    return {sum3(a.a[0], x.x, a.a[3], x.y, a.a[6], x.z),
            sum3(a.a[1], x.x, a.a[4], x.y, a.a[7], x.z),
            sum3(a.a[2], x.x, a.a[5], x.y, a.a[8], x.z)};
}
// Aᵀ x without forming Aᵀ
template <typename T> EBD V3<T> tmul(const M3<T>& a, V3<T> x) {
    return {sum3(a.a[0], x.x, a.a[1], x.y, a.a[2], x.z),
            sum3(a.a[3], x.x, a.a[4], x.y, a.a[5], x.z),
            sum3(a.a[6], x.x, a.a[7], x.y, a.a[8], x.z)};
}
template <typename T> EBD M3<T> tr(const M3<T>& a) {
    M3<T> r;
    r.a[0] = a.a[0]; r.a[1] = a.a[3]; r.a[2] = a.a[6];
    r.a[3] = a.a[1]; r.a[4] = a.a[4]; r.a[5] = a.a[7];
    r.a[6] = a.a[2]; r.a[7] = a.a[5]; r.a[8] = a.a[8];
    return r;
}
template <typename T> EBD T dot(V3<T> a, V3<T> b) { return sum3(a.x, b.x, a.y, b.y, a.z, b.z); }
template <typename T> EBD T norm(V3<T> a) { return sqrt(dot(a, a)); }
template <typename T> EBD V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> EBD M3<T> outer(V3<T> a, V3<T> b) {
    return rows(a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z);
}
template <typename T> EBD M3<T> skew(V3<T> v) { return rows(T(0), -v.z, v.y, v.z, T(0), -v.x, -v.y, v.x, T(0)); }
template <typename T> EBD V3<T> col(const M3<T>& a, int j) { return {a.a[3 * j], a.a[3 * j + 1], a.a[3 * j + 2]}; }
template <typename T> EBD M3<T> from_cols(V3<T> a, V3<T> b, V3<T> c) {
    M3<T> r; r.a[0] = a.x; r.a[1] = a.y; r.a[2] = a.z; r.a[3] = b.x; r.a[4] = b.y; r.a[5] = b.z;
    r.a[6] = c.x; r.a[7] = c.y; r.a[8] = c.z; return r;
}
template <typename T, typename S> EBD V3<T> cvt(V3<S> v) { return {T(v.x), T(v.y), T(v.z)}; }
template <typename T, typename S> EBD M3<T> cvt(const M3<S>& m) { M3<T> r; for (int k = 0; k < 9; ++k) r.a[k] = T(m.a[k]); return r; }

// ---- rotation utilities (FP64 always: src/beams/utils_beams.jl) ------------------------------

constexpr double EPS64 = 2.220446049250313e-16;

// rotation_matrix, utils_beams.jl:11-32
EBD M3<double> rotation_matrix(V3<double> th) {
    double n = norm(th);
    if (n > 10 * EPS64) {
        double s, c;
        sincos(n, &s, &c);
        M3<double> S = skew(th);
        return (eye<double>() + (s / n) * S) + (((1 - c) / (n * n)) * S) * S;
    }
    return eye<double>();
}

// toangle, utils_beams.jl:90-104
EBD V3<double> toangle(const M3<double>& R) {
    double t = R.a[0] + R.a[4] + R.a[8];
    double nv;
    if (fabs((t - 1) / 2) < 1) {
        nv = acos((t - 1) / 2);
        if (nv < 10 * EPS64) nv = 10 * EPS64;
    } else {
        nv = 10 * EPS64;
    }
    double c = 1 / (2 * sin(nv));
    V3<double> a = {at(R, 2, 1) - at(R, 1, 2), at(R, 0, 2) - at(R, 2, 0), at(R, 1, 0) - at(R, 0, 1)};
    return nv * (c * a);
}

// Tₛ⁻¹, utils_beams.jl:107-122
EBD M3<double> Ts_inv(V3<double> th) {
    double n = norm(th);
    if (n < 10 * EPS64) return eye<double>();
    M3<double> S = skew(th);
    double s2, c2;
    sincos(n / 2, &s2, &c2);
    double coef = (s2 - n / 2 * c2) / (s2 * (n * n));
    return (eye<double>() - S / 2.0) + (coef * S) * S;
}

// compute_η_μ, utils_beams.jl:180-198
EBD void eta_mu(V3<double> tb, double& eta, double& mu) {
    double t = norm(tb);
    if (t < 10 * EPS64) { eta = 1.0 / 12; mu = 1.0 / 360; return; }
    double s, c;
    sincos(t, &s, &c);
    double s2 = sin(t / 2);
    eta = (2 * s - t * (1 + c)) / ((2 * (t * t)) * s);
    double t4 = (t * t) * (t * t);
    mu = (t * (t + s) - 8 * (s2 * s2)) / ((4 * t4) * (s2 * s2));
}

// compute_K̄ₕ, utils_beams.jl:201-215
EBD M3<double> Kh_bar(V3<double> tb, V3<double> Mb, const M3<double>& Tinv, double eta, double mu) {
    M3<double> TM = outer(tb, Mb);
    M3<double> MT = tr(TM);
    M3<double> St = skew(tb), SM = skew(Mb);
    M3<double> a = (TM - 2.0 * MT) + dot(tb, Mb) * eye<double>();
    M3<double> b = (St * St) * MT;
    return ((eta * a + mu * b) - SM / 2.0) * Tinv;
}

// local_Rₑ_and_aux, utils_beams.jl:35-87 (FP64: it feeds the rotation extraction)
struct LocalFrame {
    M3<double> Re;
    V3<double> v1;
    double eta, eta11, eta12, eta21, eta22, ln;
};
EBD LocalFrame local_frame(V3<double> x1, V3<double> x2, const M3<double>& R1, const M3<double>& R2,
                           V3<double> Re0E2, double ln) {
    LocalFrame f;
    V3<double> v1 = (x2 - x1) / ln;
    V3<double> p1 = R1 * Re0E2, p2 = R2 * Re0E2;
    V3<double> p = (p1 + p2) / 2.0;
    V3<double> w3 = cross(v1, p);
    w3 = w3 / norm(w3);
    V3<double> w2 = cross(w3, v1);
    f.Re = from_cols(v1, w2, w3);
    f.v1 = v1;
    double q1 = dot(v1, p), q2 = dot(w2, p);
    double q11 = dot(v1, p1), q12 = dot(w2, p1), q21 = dot(v1, p2), q22 = dot(w2, p2);
    f.eta = q1 / q2; f.eta11 = q11 / q2; f.eta12 = q12 / q2; f.eta21 = q21 / q2; f.eta22 = q22 / q2;
    f.ln = ln;
    return f;
}

}  // namespace ebd
