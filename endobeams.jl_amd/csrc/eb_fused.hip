// eb_fused.hip — the one-workgroup predictor and corrector iteration (eb_fused.h) in a translation
// unit of its own, compiled with the library's common flags (IEEE semantics for the solve's
// finiteness and breakdown tests)
#define EB_BODIES_ONLY   // eb_api.hip defines the kernels of eb_linalg.h
#include "eb_fused.h"
#include "eb_fused_launch.h"

namespace ebd {

template <typename S> hipError_t launch_fused_predictor(const FusedArgs<S>& a, hipStream_t stream) {
    hipLaunchKernelGGL((fused_predictor<S>), dim3(1), dim3(FUSED_THREADS), 0, stream, a);
    return hipGetLastError();
}
template <typename S> hipError_t launch_fused_iteration(const FusedArgs<S>& a, bool contact, hipStream_t stream) {
    if (contact) hipLaunchKernelGGL((fused_iteration<S, true>), dim3(1), dim3(FUSED_THREADS), 0, stream, a);
    else hipLaunchKernelGGL((fused_iteration<S, false>), dim3(1), dim3(FUSED_THREADS), 0, stream, a);
    return hipGetLastError();
}

template hipError_t launch_fused_predictor<float>(const FusedArgs<float>&, hipStream_t);
template hipError_t launch_fused_predictor<double>(const FusedArgs<double>&, hipStream_t);
template hipError_t launch_fused_iteration<float>(const FusedArgs<float>&, bool, hipStream_t);
template hipError_t launch_fused_iteration<double>(const FusedArgs<double>&, bool, hipStream_t);

}  // namespace ebd
