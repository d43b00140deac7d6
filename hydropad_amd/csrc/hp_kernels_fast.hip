// hp_kernels_fast.hip -- the throughput sweeps (hp_fast.h, FP contraction on):
// FP64 (HP_FP64) and mixed FP32 reconstruction / FP64 Newton + update (HP_MIXED).
#include "hp_sweep.h"

namespace hp {

hipError_t launch_sweep_fast(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    return L.precision == PREC_MIXED ? launch_sweep_variants<FastMath, float>(b, L, s)
                                     : launch_sweep_variants<FastMath, double>(b, L, s);
}

}  // namespace hp
