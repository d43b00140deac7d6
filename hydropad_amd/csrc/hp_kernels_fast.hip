// hp_kernels_fast.hip -- the throughput sweeps (hp_fast.h, FP contraction on):
// FP64 (HP_FP64) and mixed FP32 reconstruction / FP64 Newton + update (HP_MIXED).
#include "hp_sweep.h"

namespace hp {

hipError_t launch_sweep_fast(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    if (L.g53)
        return L.precision == PREC_MIXED ? launch_sweep_variants<FastMath53, float>(b, L, s)
                                         : launch_sweep_variants<FastMath53, double>(b, L, s);
    return L.precision == PREC_MIXED ? launch_sweep_variants<FastMath, float>(b, L, s)
                                     : launch_sweep_variants<FastMath, double>(b, L, s);
}

bool params_g53(const Params& P)
{
    return P.gamma == Gamma53::gamma && P.gamma1 == Gamma53::gamma1 && P.rgamma == Gamma53::rgamma &&
           P.rgamma1 == Gamma53::rgamma1 && P.m == Gamma53::m && P.rm == Gamma53::rm && P.gf == Gamma53::gf;
}

}  // namespace hp
