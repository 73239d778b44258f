// eb_mlp.h — the network of the learned Newton warm start as seen by the solver library
// (inference on the solver's stream); the C ABI of the network is in include/endobeams_hip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct eb_mlp;

namespace ebm {
// forward pass of `rows` rows of X (device, row-major) on `stream`; *Y points at the network's
// output buffer (rows × output width), valid until the next call on this network
int mlp_infer(eb_mlp* mlp, const float* X_dev, int64_t rows, hipStream_t stream, const float** Y_dev);
// the inference workspace for `rows` rows (layer outputs only), allocated now rather than at the
// first inference
int mlp_reserve(eb_mlp* mlp, int64_t rows);
// which = 0: input width, otherwise output width
int mlp_width(const eb_mlp* mlp, int which);
}  // namespace ebm
