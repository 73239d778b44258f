// eb_element.h — launches of the element and contact kernels (eb_beam_kernel.h,
// eb_contact_kernel.h). They are compiled in a translation unit of their own, eb_element.hip, with
// structurally zero terms folded away (the flags and why the results stay the same are given
// there); the rest of the library is compiled with IEEE semantics throughout.
#pragma once
#include "eb_contact_kernel.h"

namespace ebd {

// each launches its kernel with `grid` workgroups on `stream` (errors: hipGetLastError())
template <typename P>
void launch_beam_kernel(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, BeamStage<P> st, GaussPts gp,
                              const uint8_t* active, const int32_t* beam_inst);
template <typename P>
void launch_beam_detect_kernel(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, BeamStage<P> st, GaussPts gp,
                                     const uint8_t* active, const int32_t* beam_inst, ContactArgs<P> ca,
                                     const uint8_t* is_contact, uint64_t* list, int32_t* count);
template <typename P>
void launch_contact_detect(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, GaussPts gp, ContactArgs<P> ca,
                                 uint64_t* list, int32_t* count);
template <typename P>
void launch_contact_kernel(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, BeamStage<P> st, GaussPts gp,
                                 ContactArgs<P> ca, const uint64_t* list, const int32_t* count);

}  // namespace ebd
