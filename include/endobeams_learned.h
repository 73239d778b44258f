/* endobeams_learned.h — learned Newton warm start for libendobeams_hip.so.
 *
 * No counterpart in the reference. The reference's predictor (src/predictor.jl) starts every
 * Newton iteration sequence from one linearised step. This adds an alternative start: a small
 * fully connected network predicts each node's increment over the step from the node's last
 * three converged increments and its contact force at step n. The predictor evaluates both
 * starts (the network's and the reference predictor's) and keeps, per instance, the one with the
 * smaller corrector residual, so the reference predictor stays the fallback and Newton converges
 * to the same tolerance: the network can only change the number of iterations.
 *
 * The network is trained on the device by the library's own kernels (forward, back-propagation,
 * Adam), on samples the solver records from its own converged steps (eb_learned_sample). */
#ifndef ENDOBEAMS_LEARNED_H
#define ENDOBEAMS_LEARNED_H

#include <stdint.h>

#include "endobeams_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- network: tanh MLP, FP32, widths[0] inputs ... widths[n-1] outputs (each ≤ 128) ---------
 * parameters: per layer W (out×in, row-major) then b (out), layer after layer */
typedef struct eb_mlp eb_mlp;
int eb_mlp_create(const int32_t* widths, int32_t n_widths, uint64_t seed, eb_mlp** out);
int eb_mlp_destroy(eb_mlp* mlp);
const char* eb_mlp_last_error(const eb_mlp* mlp);
int eb_mlp_num_params(const eb_mlp* mlp, int64_t* n);
int eb_mlp_get_params(eb_mlp* mlp, float* out, int64_t n);
int eb_mlp_set_params(eb_mlp* mlp, const float* in, int64_t n);   /* also resets the Adam moments */
/* Y = f(X) for `rows` rows (host buffers, row-major) */
int eb_mlp_forward(eb_mlp* mlp, const float* X, int64_t rows, float* Y);
/* mean squared error over rows×outputs and its gradient with respect to every parameter */
int eb_mlp_loss_gradient(eb_mlp* mlp, const float* X, const float* T, int64_t rows, float* loss, float* grad);
/* training set (copied to the device) and Adam epochs over it in shuffled minibatches;
 * loss_per_epoch (epochs entries, may be NULL) receives the mean minibatch loss of each epoch */
int eb_mlp_set_dataset(eb_mlp* mlp, const float* X, const float* T, int64_t rows);
int eb_mlp_train(eb_mlp* mlp, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss_per_epoch);

/* --- solver side -----------------------------------------------------------------------------
 * Per node i (inputs, EB_LEARNED_INPUTS): p — the reference predictor's increment of this step
 * (6), then r1, r2, r3 — the node's last three converged increments [Δu; w] (6 each, newest first)
 * expressed at the current Δt (each recorded as a rate over its own step), then the node's contact
 * force Tᶜⁿ (translational part, 3). Output (EB_LEARNED_OUTPUTS): the correction c to the
 * reference predictor's start, ΔD_i = p + c on free dofs. Inputs are multiplied by in_scale and
 * outputs by out_scale. */
#define EB_LEARNED_INPUTS 27
#define EB_LEARNED_OUTPUTS 6
/* samples of the step just converged (call after the corrector, before eb_update_converged):
 * X (n_nodes × 27) unscaled inputs as the predictor saw them, T (n_nodes × 6) the unscaled target
 * correction ΔD* - p. Fails until three steps have converged since eb_initialize_state_simulation. */
int eb_learned_sample(eb_ctx* ctx, float* X, float* T);
/* installs the network (NULL: reference predictor only); the context keeps the pointer and
 * allocates here its own buffers for the network's layer outputs over its nodes (device memory,
 * replacing those of a network installed before). Several contexts may use one network, also
 * from different host threads; it must not be trained meanwhile. */
int eb_set_learned_predictor(eb_ctx* ctx, eb_mlp* mlp, const float* in_scale, const float* out_scale);
/* A network whose start keeps losing is not evaluated at every step: after an evaluation in
 * which no instance kept the network's start, the next 1, 2, 4, ... (at most max_skip)
 * predictions use the reference predictor alone; an evaluation in which some instance keeps it
 * resets the wait. max_skip = 0 evaluates the network at every prediction; the default is 16. */
int eb_set_learned_backoff(eb_ctx* ctx, int32_t max_skip);
/* instance-predictions that kept the network's start [0], that fell back to the reference's after
 * evaluating both [1], and predictions that used the reference's without evaluating the network
 * under the backoff [2] */
int eb_get_predictor_stats(const eb_ctx* ctx, int64_t* out3);
/* for checks: the increment ΔD of the network's start at the last prediction that evaluated the
 * network (n_dofs doubles: p + out_scale ⊙ Y on free dofs, p on fixed dofs) */
int eb_get_learned_increment(eb_ctx* ctx, double* out, int64_t count);

#ifdef __cplusplus
}
#endif

#endif /* ENDOBEAMS_LEARNED_H */
