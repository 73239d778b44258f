// eb_element.hip — the element and contact kernels, in a translation unit compiled with
// -fno-signed-zeros -ffinite-math-only (on top of the library's -O3 -ffp-contract=off).
//
// The element code evaluates the reference's expressions with dense 3x3 products, as the
// reference's StaticArrays code does, and many operands are structurally zero: skew matrices,
// the sparse G, P, H and A blocks (compute_beams.jl:233-421), diagonal matrices. Under IEEE rules
// the compiler must still multiply by those zeros (0·x is NaN for x = ±∞, and x + 0 may flip the
// sign of a zero). With these two flags it folds 0·x to 0 and x + 0 to x; every product with no
// constant-zero operand is still computed, in the same order (no reassociation, no contraction),
// so every finite result is the one the unfolded code gives, up to the sign of results that are
// exactly zero. No element or contact function tests for NaN or infinity. The fold removes 20-35 %
// of the VALU instructions and a quarter of the registers of the element kernel's wavefronts.
#include "eb_element.h"

namespace ebd {

// The precision form's element kernel as two launches of two parts each (throughput: the tiles
// fill the device several times over) from this many tiles on; below, one launch of all four parts
// (latency: the parts of a tile run side by side, which is what a small structure's few tiles need).
// Measured per element assembly on batches of the 4 800-beam stent (four-part / two launches):
// 1 200 tiles 101 / 114 µs, 2 400 tiles 170 / 169, 4 800 tiles 297 / 285, 9 600 tiles 540 / 485.
constexpr unsigned SPLIT_MIN_TILES = 2048;

template <typename P>
void launch_beam_kernel(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, BeamStage<P> st,
                              GaussPts gp, const uint8_t* active, const int32_t* beam_inst) {
    if constexpr (sizeof(P) == sizeof(float)) {
        if (grid >= SPLIT_MIN_TILES) {
            hipLaunchKernelGGL((beam_kernel_half<P, PART_STATIC>), dim3(grid), dim3(2 * BEAM_TILE), 0, stream, ns, bc, st,
                               gp, active, beam_inst);
            hipLaunchKernelGGL((beam_kernel_half<P, PART_C13>), dim3(grid), dim3(2 * BEAM_TILE), 0, stream, ns, bc, st,
                               gp, active, beam_inst);
            return;
        }
    }
    hipLaunchKernelGGL(beam_kernel<P>, dim3(grid), dim3(4 * BEAM_TILE), 0, stream, ns, bc, st, gp, active, beam_inst);
}

template <typename P>
void launch_beam_detect_kernel(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, BeamStage<P> st,
                                     GaussPts gp, const uint8_t* active, const int32_t* beam_inst, ContactArgs<P> ca,
                                     const uint8_t* is_contact, uint64_t* list, int32_t* count) {
    if constexpr (sizeof(P) == sizeof(float)) {
        if (grid >= SPLIT_MIN_TILES) {
            hipLaunchKernelGGL(beam_detect_kernel_half<P>, dim3(grid), dim3(2 * BEAM_TILE), 0, stream, ns, bc, st, gp,
                               active, beam_inst, ca, is_contact, list, count);
            hipLaunchKernelGGL((beam_kernel_half<P, PART_C13>), dim3(grid), dim3(2 * BEAM_TILE), 0, stream, ns, bc, st,
                               gp, active, beam_inst);
            return;
        }
    }
    hipLaunchKernelGGL(beam_detect_kernel<P>, dim3(grid), dim3(4 * BEAM_TILE), 0, stream, ns, bc, st, gp, active,
                       beam_inst, ca, is_contact, list, count);
}

template <typename P>
void launch_contact_detect(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, GaussPts gp,
                                 ContactArgs<P> ca, uint64_t* list, int32_t* count) {
    hipLaunchKernelGGL(contact_detect<P>, dim3(grid), dim3(256), 0, stream, ns, bc, gp, ca, list, count);
}

template <typename P>
void launch_contact_kernel(unsigned grid, hipStream_t stream, NodeState ns, BeamConsts bc, BeamStage<P> st,
                                 GaussPts gp, ContactArgs<P> ca, const uint64_t* list, const int32_t* count) {
    if constexpr (sizeof(P) == sizeof(float)) {
        // the FP64 contact forces first, one lane per listed beam (grid * CONTACT_TILE lanes cover
        // the list), then the FP32 tangent, four wavefronts per 64 beams
        hipLaunchKernelGGL(contact_force_kernel<P>, dim3((unsigned)((grid * (int64_t)CONTACT_TILE + 255) / 256)), dim3(256),
                           0, stream, ns, bc, st, gp, ca, list, count);
    }
    hipLaunchKernelGGL(contact_kernel<P>, dim3(grid), dim3(4 * CONTACT_TILE), 0, stream, ns, bc, st, gp, ca, list,
                       count);
}

#define EB_ELEMENT_INSTANTIATE(P)                                                                                   \
    template void launch_beam_kernel<P>(unsigned, hipStream_t, NodeState, BeamConsts, BeamStage<P>, GaussPts, \
                                              const uint8_t*, const int32_t*);                                      \
    template void launch_beam_detect_kernel<P>(unsigned, hipStream_t, NodeState, BeamConsts, BeamStage<P>,    \
                                                     GaussPts, const uint8_t*, const int32_t*, ContactArgs<P>,      \
                                                     const uint8_t*, uint64_t*, int32_t*);                          \
    template void launch_contact_detect<P>(unsigned, hipStream_t, NodeState, BeamConsts, GaussPts,            \
                                                 ContactArgs<P>, uint64_t*, int32_t*);                              \
    template void launch_contact_kernel<P>(unsigned, hipStream_t, NodeState, BeamConsts, BeamStage<P>,        \
                                                 GaussPts, ContactArgs<P>, const uint64_t*, const int32_t*);
EB_ELEMENT_INSTANTIATE(float)
EB_ELEMENT_INSTANTIATE(double)

}  // namespace ebd
