// eb_learned.h — device side of the learned Newton warm start: per-node history of converged
// increments, network inputs/targets and the application of a predicted increment.
#pragma once
#include "eb_dmath.h"
#include "eb_linalg.h"

namespace ebd {

constexpr int LEARN_IN = 27, LEARN_OUT = 6;

// Per-node history of the last three converged steps of each instance: three slots of 6N rates
// ([u - uⁿ; w] / Δt), slot k at hist + 6N k; newest[s] is the slot instance s wrote last, the two
// before it are (newest + 2) % 3 and (newest + 1) % 3.

// rate of the step just converged, [u - uⁿ; w] / Δt with the step's own Δt, into the slot
// newest[s] of the instances selected in `sel` (all when null)
__global__ void __launch_bounds__(256) learned_record(int64_t N, const double* u, const double* u_n, const double* w,
                                                      const StepCoef* __restrict__ coef, const int32_t* node_inst,
                                                      const uint8_t* sel, const int32_t* newest, double* hist) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int32_t s = node_inst[n];
    if (sel && !sel[s]) return;
    const double inv_dt = 1.0 / coef[s].dt;
    double* rate = hist + 6 * N * newest[s];
    for (int c = 0; c < 3; ++c) {
        rate[6 * n + c] = (u[3 * n + c] - u_n[3 * n + c]) * inv_dt;
        rate[6 * n + 3 + c] = w[3 * n + c] * inv_dt;
    }
}

// newest[s] = (newest[s] + 1) % 3 for the instances selected in `sel` (all when null): the slot the
// step just converged is recorded in
__global__ void __launch_bounds__(256) learned_advance(int64_t S, const uint8_t* sel, int32_t* newest) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || (sel && !sel[s])) return;
    newest[s] = (newest[s] + 1) % 3;
}

// inputs per node: the reference predictor's increment p, the three recorded rates × Δt (newest
// first), then Tᶜⁿ (translations), × scale
__global__ void __launch_bounds__(256) learned_features(int64_t N, const double* pdD, const double* hist,
                                                        const int32_t* newest, const int32_t* node_inst,
                                                        const double* Tc_n, const StepCoef* __restrict__ coef,
                                                        const float* scale, float* X) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int32_t s = node_inst[n];
    const double dt = coef[s].dt;
    const int32_t h0 = newest[s];
    float* x = X + LEARN_IN * n;
    for (int c = 0; c < 6; ++c) x[c] = float(pdD[6 * n + c]) * scale[c];
    for (int k = 0; k < 3; ++k) {
        const double* h = hist + 6 * N * ((h0 + 3 - k) % 3);
        for (int c = 0; c < 6; ++c) x[6 + 6 * k + c] = float(h[6 * n + c] * dt) * scale[6 + 6 * k + c];
    }
    for (int c = 0; c < 3; ++c) x[24 + c] = float(Tc_n[6 * n + c]) * scale[24 + c];
}

// target correction of the step just converged: [u - uⁿ; w] - p, × scale
__global__ void __launch_bounds__(256) learned_targets(int64_t N, const double* u, const double* u_n, const double* w,
                                                       const double* pdD, const float* scale, float* T) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int c = 0; c < 3; ++c) {
        T[LEARN_OUT * n + c] = float((u[3 * n + c] - u_n[3 * n + c]) - pdD[6 * n + c]) * scale[c];
        T[LEARN_OUT * n + 3 + c] = float(w[3 * n + c] - pdD[6 * n + 3 + c]) * scale[3 + c];
    }
}

// ΔD = p + out_scale ⊙ Y on free dofs, p on fixed dofs (their imposed increment)
__global__ void __launch_bounds__(256) learned_apply(int64_t N, const float* Y, const double* pdD,
                                                    const float* out_scale, const uint8_t* fixed, double* dD) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * N) return;
    const int64_t n = i / 6;
    const int c = (int)(i % 6);
    dD[i] = fixed[i] ? pdD[i] : pdD[i] + double(Y[LEARN_OUT * n + c]) * double(out_scale[c]);
}

// current node state ← snapshot for the nodes of instances whose flag is set (or all when
// flag is null); count doubles per node
__global__ void __launch_bounds__(256) copy_node_fields(int64_t N, int per_node, const double* src, double* dst,
                                                        const uint8_t* flag, const int32_t* node_inst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * per_node) return;
    if (flag && !flag[node_inst[i / per_node]]) return;
    dst[i] = src[i];
}

// one workgroup: per eligible instance (active in the step, three converged steps recorded)
// the start with the smaller relative residual (the network's when it is not larger, and
// finite); counts[0], counts[1]: running totals of the two choices over eligible instances,
// counts[2]: instances taking the network's start in this call
__global__ void __launch_bounds__(256) learned_choose(int ninst, const double* rn_plain, const double* rn_learned,
                                                      const uint8_t* eligible, const uint8_t* active, uint8_t* use_plain,
                                                      uint8_t* use_learned, int32_t* counts) {
    __shared__ int kept, seen;
    if (threadIdx.x == 0) { kept = 0; seen = 0; }
    __syncthreads();
    int mine = 0, mine_seen = 0;
    for (int s = threadIdx.x; s < ninst; s += blockDim.x) {
        const double a = rn_plain[2 * s], b = rn_learned[2 * s];
        const bool el = eligible[s] && active[s];
        const bool learned = el && b <= a && b == b;
        use_plain[s] = learned ? 0 : 1;
        use_learned[s] = learned ? 1 : 0;
        mine += learned ? 1 : 0;
        mine_seen += el ? 1 : 0;
    }
    atomicAdd(&kept, mine);
    atomicAdd(&seen, mine_seen);
    __syncthreads();
    if (threadIdx.x == 0) {
        counts[0] += kept;
        counts[1] += seen - kept;
        counts[2] = kept;
    }
}

}  // namespace ebd
