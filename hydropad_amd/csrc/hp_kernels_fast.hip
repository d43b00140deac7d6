// hp_kernels_fast.hip -- the throughput sweeps (hp_fast.h, FP contraction on):
// FP64 (HP_FP64) and mixed FP32 reconstruction / FP64 Newton + update (HP_MIXED).
#include "hp_sweep.h"

namespace hp {

template <class R>
static hipError_t launch_fast_r(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    if (L.grav) return L.cosmo ? launch_sweep_t<FastMath, R, true, true>(b, L, s)
                               : launch_sweep_t<FastMath, R, true, false>(b, L, s);
    return L.cosmo ? launch_sweep_t<FastMath, R, false, true>(b, L, s)
                   : launch_sweep_t<FastMath, R, false, false>(b, L, s);
}

hipError_t launch_sweep_fast(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    return L.precision == PREC_MIXED ? launch_fast_r<float>(b, L, s) : launch_fast_r<double>(b, L, s);
}

}  // namespace hp
