// eb_output.h — the per-beam output points of write_VTK (recompute_at_gausspts_beams!,
// src/IO/visualization_beams.jl:132-181): position with the cubic bending offset, displacement,
// velocity, strain and stress at `np` points evenly spaced along each beam, from the current node
// state. One lane per beam: the element frame is computed once and the points written after it.
#pragma once
#include "eb_beam_kernel.h"

namespace ebd {

__global__ void __launch_bounds__(256) beam_output_points(NodeState ns, BeamConsts bc, int np, double* pos, double* disp,
                                                          double* vel, double* strain, double* stress) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= bc.E) return;
    const int64_t E = bc.E;
    const int32_t a = bc.n1[e], b = bc.n2[e];
    const V3<double> u1 = ldv(ns.u, a), u2 = ldv(ns.u, b);
    const V3<double> x1 = ldv(ns.X0, a) + u1, x2 = ldv(ns.X0, b) + u2;
    const V3<double> v1 = ldv(ns.ud, a), v2 = ldv(ns.ud, b);
    const double ln = norm(x2 - x1);
    const double l0 = bc.l0[e];
    M3<double> Re0;
    for (int k = 0; k < 9; ++k) Re0.a[k] = bc.Re0[k * E + e];
    const M3<double> R1 = ldm(ns.R, a), R2 = ldm(ns.R, b);
    const LocalFrame f = local_frame(x1, x2, R1, R2, col(Re0, 1), ln);
    const M3<double> ReT = tr(f.Re);
    const V3<double> Tb1 = toangle((ReT * R1) * Re0), Tb2 = toangle((ReT * R2) * Re0);
    const double eps = 1 - ln / l0;
    const double Emod = bc.props[1 * E + e];
    for (int i = 0; i < np; ++i) {
        const int64_t k = e * np + i;
        const double z = np > 1 ? -1.0 + (2.0 * i) / (np - 1) : -1.0;   // range(-1, 1, length = np)
        const double xi = l0 * (z + 1) / 2;
        const double N1 = 1 - xi / l0;
        const double N2 = 1 - N1;
        const double c = 1 - xi / l0;
        const double N3 = xi * (c * c);
        const double N4 = -(1 - xi / l0) * ((xi * xi) / l0);
        const V3<double> ut = {0.0, N3 * Tb1.z + N4 * Tb2.z, -N3 * Tb1.y - N4 * Tb2.y};
        const V3<double> xg = (N1 * x1 + N2 * x2) + f.Re * ut;
        if (pos) { pos[3 * k] = xg.x; pos[3 * k + 1] = xg.y; pos[3 * k + 2] = xg.z; }
        if (disp) {
            const V3<double> d = N1 * u1 + N2 * u2;
            disp[3 * k] = d.x; disp[3 * k + 1] = d.y; disp[3 * k + 2] = d.z;
        }
        if (vel) {
            const V3<double> w = N1 * v1 + N2 * v2;
            vel[3 * k] = w.x; vel[3 * k + 1] = w.y; vel[3 * k + 2] = w.z;
        }
        if (strain) strain[k] = eps;
        if (stress) stress[k] = Emod * eps;
    }
}

}  // namespace ebd
