// eb_fused_launch.h — launchers of the one-workgroup predictor and corrector iteration (eb_fused.h,
// compiled in eb_fused.hip)
#pragma once
#include "eb_fused.h"

namespace ebd {
template <typename S> hipError_t launch_fused_predictor(const FusedArgs<S>& a, hipStream_t stream);
template <typename S> hipError_t launch_fused_iteration(const FusedArgs<S>& a, bool contact, hipStream_t stream);
}
