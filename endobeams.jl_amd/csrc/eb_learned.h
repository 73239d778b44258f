// eb_learned.h — device side of the learned Newton warm start: per-node history of converged
// increments, network inputs/targets and the application of a predicted increment.
#pragma once
#include "eb_dmath.h"

namespace ebd {

constexpr int LEARN_IN = 27, LEARN_OUT = 6;

// rate of the step just converged: [u - uⁿ; w] / Δt per node
__global__ void __launch_bounds__(256) learned_record(int64_t N, const double* u, const double* u_n, const double* w,
                                                      double inv_dt, double* rate) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int c = 0; c < 3; ++c) {
        rate[6 * n + c] = (u[3 * n + c] - u_n[3 * n + c]) * inv_dt;
        rate[6 * n + 3 + c] = w[3 * n + c] * inv_dt;
    }
}

// inputs per node: the reference predictor's increment p, the three recorded rates × Δt (newest
// first), then Tᶜⁿ (translations), × scale
__global__ void __launch_bounds__(256) learned_features(int64_t N, const double* pdD, const double* h0,
                                                        const double* h1, const double* h2, const double* Tc_n,
                                                        double dt, const float* scale, float* X) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float* x = X + LEARN_IN * n;
    for (int c = 0; c < 6; ++c) x[c] = float(pdD[6 * n + c]) * scale[c];
    const double* h[3] = {h0, h1, h2};
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 6; ++c) x[6 + 6 * k + c] = float(h[k][6 * n + c] * dt) * scale[6 + 6 * k + c];
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

// one workgroup: per instance the start with the smaller relative residual (the network's when
// it is not larger, and finite); counts[0], counts[1]: running totals of the two choices,
// counts[2]: instances taking the network's start in this call
__global__ void __launch_bounds__(256) learned_choose(int ninst, const double* rn_plain, const double* rn_learned,
                                                      uint8_t* use_plain, uint8_t* use_learned, int32_t* counts) {
    __shared__ int kept;
    if (threadIdx.x == 0) kept = 0;
    __syncthreads();
    int mine = 0;
    for (int s = threadIdx.x; s < ninst; s += blockDim.x) {
        const double a = rn_plain[2 * s], b = rn_learned[2 * s];
        const bool learned = b <= a && b == b;
        use_plain[s] = learned ? 0 : 1;
        use_learned[s] = learned ? 1 : 0;
        mine += learned ? 1 : 0;
    }
    atomicAdd(&kept, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        counts[0] += kept;
        counts[1] += ninst - kept;
        counts[2] = kept;
    }
}

}  // namespace ebd
