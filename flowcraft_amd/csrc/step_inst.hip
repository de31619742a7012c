// Explicit instantiations of the step-kernel launcher for one lattice and storage
// format, selected at compile time: -DFC_Q=19|15 -DFC_STORE=StoreF32|StoreF16|StoreBF16|StoreF64|
// StoreF16SR|StoreBF16SR.
#include "step_launch.hpp"

#ifndef FC_Q
#error "compile with -DFC_Q=<15|19>"
#endif
#ifndef FC_STORE
#error "compile with -DFC_STORE=<StoreF32|StoreF16|StoreBF16|StoreF64>"
#endif

namespace fc {

template <int Q, class S, class Lay>
hipError_t LaunchStep(const StepArgs& a, unsigned tiles, int cpt, bool periodic, bool resident, hipStream_t stream) {
    const dim3 grid(tiles), block(StepThreads(sizeof(typename S::T)));
    if (periodic) {
        if (cpt == 2)
            hipLaunchKernelGGL((KStep<Q, S, Lay, 2, true>), grid, block, 0, stream, a);
        else if (resident)
            hipLaunchKernelGGL((KStep<Q, S, Lay, 1, true, true>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((KStep<Q, S, Lay, 1, true>), grid, block, 0, stream, a);
    } else {
        if (cpt == 2)
            hipLaunchKernelGGL((KStep<Q, S, Lay, 2, false>), grid, block, 0, stream, a);
        else if (resident)
            hipLaunchKernelGGL((KStep<Q, S, Lay, 1, false, true>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((KStep<Q, S, Lay, 1, false>), grid, block, 0, stream, a);
    }
    return hipGetLastError();
}

template hipError_t LaunchStep<FC_Q, FC_STORE<FC_Q>, PingPong>(const StepArgs&, unsigned, int, bool, bool, hipStream_t);
template hipError_t LaunchStep<FC_Q, FC_STORE<FC_Q>, InPlace>(const StepArgs&, unsigned, int, bool, bool, hipStream_t);

}  // namespace fc
