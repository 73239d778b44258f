// flopcount.cpp — TEST INFRASTRUCTURE: operation-counting build of the oracle's element kernels.
//
// Instantiates eb_elements.h with a scalar that counts every +, -, ×, ÷ (FP64 arithmetic of the
// restated expression DAG, so also products with the structural zeros that dense 3x3 algebra
// carries) and, separately, sqrt and the transcendentals. It also counts the operations left
// once exact zeros are folded (x·0, x ± 0, 0 / x, the work of a kernel compiled to drop them): on
// a generic element state every exact zero is a structural one. That count is the algorithmic
// flop figure the benchmark's roofline divides by the element kernel's launch time.
#include <cmath>
#include <cstdint>
#include <type_traits>

namespace flopcount {
struct Counts { long long flops = 0, sqrt = 0, trans = 0, nonzero = 0; };
static Counts g;

struct C {
    double v = 0;
    C() = default;
    template <typename T, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    C(T x) : v((double)x) {}
    explicit operator double() const { return v; }
    explicit operator int64_t() const { return (int64_t)v; }
    C operator-() const { return C(-v); }
    C& operator+=(const C& o) { ++g.flops; g.nonzero += (v != 0 && o.v != 0); v += o.v; return *this; }
};
inline void both(const C& a, const C& b) { ++g.flops; g.nonzero += (a.v != 0 && b.v != 0); }
inline C operator+(const C& a, const C& b) { both(a, b); return C(a.v + b.v); }
inline C operator-(const C& a, const C& b) { both(a, b); return C(a.v - b.v); }
inline C operator*(const C& a, const C& b) { both(a, b); return C(a.v * b.v); }
inline C operator/(const C& a, const C& b) { ++g.flops; g.nonzero += (a.v != 0); return C(a.v / b.v); }
template <typename T> inline C operator+(const C& a, T b) { return a + C(b); }
template <typename T> inline C operator+(T a, const C& b) { return C(a) + b; }
template <typename T> inline C operator-(const C& a, T b) { return a - C(b); }
template <typename T> inline C operator-(T a, const C& b) { return C(a) - b; }
template <typename T> inline C operator*(const C& a, T b) { return a * C(b); }
template <typename T> inline C operator*(T a, const C& b) { return C(a) * b; }
template <typename T> inline C operator/(const C& a, T b) { return a / C(b); }
template <typename T> inline C operator/(T a, const C& b) { return C(a) / b; }
inline bool operator<(const C& a, const C& b) { return a.v < b.v; }
inline bool operator>(const C& a, const C& b) { return a.v > b.v; }
inline bool operator<=(const C& a, const C& b) { return a.v <= b.v; }
inline bool operator>=(const C& a, const C& b) { return a.v >= b.v; }
inline bool operator==(const C& a, const C& b) { return a.v == b.v; }
inline bool operator!=(const C& a, const C& b) { return a.v != b.v; }
inline C sqrt(const C& a) { ++g.sqrt; return C(std::sqrt(a.v)); }
inline C sin(const C& a) { ++g.trans; return C(std::sin(a.v)); }
inline C cos(const C& a) { ++g.trans; return C(std::cos(a.v)); }
inline C acos(const C& a) { ++g.trans; return C(std::acos(a.v)); }
inline C fabs(const C& a) { return C(std::fabs(a.v)); }
inline C floor(const C& a) { return C(std::floor(a.v)); }
}  // namespace flopcount

#define EBO_REAL flopcount::C
#define ebo ebo_counted
#include "eb_elements.h"
#undef ebo

namespace {
using namespace ebo_counted;
ElemIn unpack(const double* x) {
    ElemIn in;
    V3* vs[12] = {&in.X1, &in.X2, &in.u1, &in.u2, &in.ud1, &in.ud2, &in.udd1, &in.udd2, &in.wd1, &in.wd2, &in.wdd1, &in.wdd2};
    for (int k = 0; k < 12; ++k) *vs[k] = v3(x[3 * k], x[3 * k + 1], x[3 * k + 2]);
    M3* ms[4] = {&in.R1, &in.R2, &in.dR1, &in.dR2};
    for (int k = 0; k < 4; ++k) for (int q = 0; q < 9; ++q) ms[k]->m[q] = x[36 + 9 * k + q];
    in.l0 = x[72];
    for (int q = 0; q < 9; ++q) in.Re0.m[q] = x[73 + q];
    return in;
}
}  // namespace

extern "C" {
// counts for one compute_beams call: out4 = {flops, sqrt, transcendentals, flops without exact zeros}
int ebo_count_beam(const double* elem82, const double* props14, int n_gauss, const double* gw, const double* gz,
                   long long* out4) {
    ElemIn in = unpack(elem82);
    flopcount::C p[14];
    for (int k = 0; k < 14; ++k) p[k] = props14[k];
    BeamProps bp = props_from(p);
    Gauss g;
    g.n = n_gauss;
    for (int k = 0; k < n_gauss; ++k) { g.w[k] = gw[k]; g.z[k] = gz[k]; }
    BeamOut o;
    flopcount::g = flopcount::Counts();
    compute_beams(in, bp, g, o);
    out4[0] = flopcount::g.flops; out4[1] = flopcount::g.sqrt; out4[2] = flopcount::g.trans;
    out4[3] = flopcount::g.nonzero;
    return 0;
}
}
