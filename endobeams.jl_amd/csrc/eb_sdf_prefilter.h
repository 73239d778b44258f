// eb_sdf_prefilter.h — quadratic B-spline prefilter of a signed distance field on the device
// (interpolate(field, BSpline(Quadratic(Reflect(OnCell())))), src/interactions/def_interactions.jl:77).
// The separable system c₀ = c₁, c_{i-1}/8 + 3c_i/4 + c_{i+1}/8 = y_i, c_{n+1} = c_n is solved
// along x, then y, then z, one lane per grid line, with the host's LU factors and the host's
// operation order, so the coefficients equal the host prefilter's bit for bit. Padded output
// (n0+2)(n1+2)(n2+2), x fastest. Lines along y and z are indexed by x, so a wavefront's accesses
// to one grid plane are contiguous.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ebd {

struct TriLU {
    const double* cp;   // n: modified super-diagonal
    const double* inv;  // n: reciprocal pivots
    int64_t n;
};

// y: n values at stride sy (may alias c + sc), c: n + 2 values at stride sc
__device__ inline void tri_solve_line(const TriLU t, const double* y, int64_t sy, double* c, int64_t sc) {
    const int64_t n = t.n;
    double prev = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double v = (y[i * sy] - (i > 0 ? 0.125 * prev : 0.0)) * t.inv[i];
        c[(i + 1) * sc] = v;
        prev = v;
    }
    for (int64_t i = n - 2; i >= 0; --i) c[(i + 1) * sc] = c[(i + 1) * sc] - t.cp[i] * c[(i + 2) * sc];
    c[0] = c[sc];
    c[(n + 1) * sc] = c[n * sc];
}

struct SdfGrid { int64_t n0, n1, n2; };

// lines along x: one per (j, k) of the samples
__global__ void __launch_bounds__(256) sdf_prefilter_x(SdfGrid g, TriLU t, const double* __restrict__ samples, double* coef) {
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= g.n1 * g.n2) return;
    const int64_t j = line % g.n1, k = line / g.n1;
    const int64_t P0 = g.n0 + 2, P1 = g.n1 + 2;
    tri_solve_line(t, samples + (k * g.n1 + j) * g.n0, 1, coef + ((k + 1) * P1 + (j + 1)) * P0, 1);
}

// lines along y, in place: one per padded x and sample plane k
__global__ void __launch_bounds__(256) sdf_prefilter_y(SdfGrid g, TriLU t, double* coef) {
    const int64_t P0 = g.n0 + 2, P1 = g.n1 + 2;
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= P0 * g.n2) return;
    const int64_t i = line % P0, k = line / P0;
    double* base = coef + (k + 1) * P1 * P0 + i;
    tri_solve_line(t, base + P0, P0, base, P0);
}

// lines along z, in place: one per padded (x, y)
__global__ void __launch_bounds__(256) sdf_prefilter_z(SdfGrid g, TriLU t, double* coef) {
    const int64_t P0 = g.n0 + 2, P1 = g.n1 + 2;
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= P0 * P1) return;
    double* base = coef + line;
    tri_solve_line(t, base + P1 * P0, P1 * P0, base, P1 * P0);
}

// coarse lower bounds: one lane per coarse cell, the minimum of the stored coefficients over the
// padded indices block·c … block·c + block + 1 of each axis (clipped to the padded grid)
template <typename P>
__global__ void __launch_bounds__(256) sdf_block_min(SdfGrid g, int block, const P* __restrict__ coef, int32_t cn0,
                                                     int32_t cn1, int32_t cn2, P* cmin) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)cn0 * cn1 * cn2) return;
    const int64_t c0 = q % cn0, c1 = (q / cn0) % cn1, c2 = q / ((int64_t)cn0 * cn1);
    const int64_t P0 = g.n0 + 2, P1 = g.n1 + 2, P2 = g.n2 + 2;
    const int64_t i0 = block * c0, j0 = block * c1, k0 = block * c2;
    const int64_t i1 = i0 + block + 1 < P0 - 1 ? i0 + block + 1 : P0 - 1;
    const int64_t j1 = j0 + block + 1 < P1 - 1 ? j0 + block + 1 : P1 - 1;
    const int64_t k1 = k0 + block + 1 < P2 - 1 ? k0 + block + 1 : P2 - 1;
    P m = coef[(k0 * P1 + j0) * P0 + i0];
    for (int64_t k = k0; k <= k1; ++k)
        for (int64_t j = j0; j <= j1; ++j)
            for (int64_t i = i0; i <= i1; ++i) {
                const P v = coef[(k * P1 + j) * P0 + i];
                m = v < m ? v : m;
            }
    cmin[q] = m;
}

template <typename P>
__global__ void __launch_bounds__(256) sdf_to_storage(int64_t n, const double* __restrict__ src, P* dst) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        dst[q] = P(src[q]);
}

}  // namespace ebd
