// eb_mlp.h — the network of the learned Newton warm start as seen by the solver library
// (inference on the solver's stream); the C ABI of the network is in include/endobeams_hip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct eb_mlp;

namespace ebm {
// forward pass of `rows` rows of X (device, row-major) on `stream`, layer l's output into the
// caller's act[l] (rows × mlp_layer_width(l)); *Y_dev = the last one. The network's parameters
// are only read: contexts may infer with one network concurrently, each with its own buffers
// and stream, as long as nobody trains it meanwhile.
int mlp_infer(eb_mlp* mlp, const float* X_dev, int64_t rows, hipStream_t stream, float* const* act,
              const float** Y_dev);
int mlp_layers(const eb_mlp* mlp);
int mlp_layer_width(const eb_mlp* mlp, int l);   // outputs of layer l
// which = 0: input width, otherwise output width
int mlp_width(const eb_mlp* mlp, int which);
}  // namespace ebm
