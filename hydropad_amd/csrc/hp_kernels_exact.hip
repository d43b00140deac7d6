// hp_kernels_exact.hip -- the FP64 sweep in the reference's operation order
// (built with -ffp-contract=off): bit-faithful to the reference except where
// OCML's pow and glibc's differ in the last bit.  Selected by HP_FP64_EXACT.
#include "hp_sweep.h"

namespace hp {

hipError_t launch_sweep_exact(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    return launch_sweep_variants<ExactMath, double>(b, L, s);
}

}  // namespace hp
