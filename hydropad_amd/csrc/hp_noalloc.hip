// hp_noalloc.hip -- one directional sweep on caller-owned device buffers with
// no device allocation anywhere on the call path (include/hydropad_gpu.h,
// hp_ppmsolver_noalloc).  The sweep kernels are hp_ppmsolver's; what a
// context would allocate besides the fields -- the device copy of the 19
// scalars and the status block (CFL maximum, Newton and clamp counters) --
// lives in the file-scope array `scratch` below, one copy per GPU for the
// life of the process.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>

#include "../../include/hydropad_gpu.h"
#include "hp_core.h"
#include "hp_kernels.h"

namespace {

// Scalars block then status block, each 256-byte aligned.
constexpr size_t PRM_OFF = 0;
constexpr size_t ST_OFF = (sizeof(hp::Params) + 255) / 256 * 256;
constexpr size_t SCRATCH_BYTES = ST_OFF + (sizeof(hp::DevStatus) + 255) / 256 * 256;
static_assert(SCRATCH_BYTES == HP_NOALLOC_SCRATCH_BYTES, "scratch size stated in the header");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_ptr[MAX_DEVICES];   // device address of `scratch` on each GPU, set once

// Device address of this GPU's copy of `scratch` (the array's name is only a
// host-side stand-in), looked up once per device with that device current.
int scratch_of_current_device(unsigned char** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return HP_E_HIP;
    if (dev < 0 || dev >= MAX_DEVICES) return HP_E_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HP_E_HIP;
        g_ptr[dev] = (unsigned char*)p;
    }
    *out = g_ptr[dev];
    return HP_OK;
}

// Write the scalars and clear the status block with a kernel (no memset or
// host copy, so the sequence can also be captured in a graph).
__global__ void scratch_init_kernel(hp::Params P, hp::Params* prm, hp::DevStatus* st, int newton_cap)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    *prm = P;
    hp::DevStatus z;
    memset(&z, 0, sizeof(z));
    z.newton_cap = newton_cap;
    *st = z;
}

}  // namespace

extern "C" int hp_ppmsolver_noalloc(const hp_params* prm, const double* p, const double* rho, const double* vx,
                                    const double* vy, const double* vz, double* p_out, double* rho_out,
                                    double* vx_out, double* vy_out, double* vz_out, const double* gforce, int nx,
                                    int ny, int nz, int nbound, int ndirection, int precision, int newton_limit,
                                    hp_counters* counters, void* stream)
{
    if (!prm || !p || !rho || !vx || !vy || !vz || !p_out || !rho_out || !vx_out || !vy_out || !vz_out)
        return HP_E_ARG;
    if (ndirection < 1 || ndirection > 3 || newton_limit < 0) return HP_E_ARG;
    if (nbound != 3 || nx < 2 * nbound + 3 || ny < 2 * nbound + 3 || nz < 2 * nbound + 3) return HP_E_ARG;
    if ((long long)nx * ny * nz >= (1LL << 31)) return HP_E_ARG;   // 32-bit cell indexing in the sweeps
    if (precision != HP_FP64 && precision != HP_MIXED && precision != HP_FP64_EXACT) return HP_E_ARG;
    if (!(prm->dx > 0.0) || !(prm->gamma > 1.0)) return HP_E_ARG;
    const double* in[5] = {p, rho, vx, vy, vz};
    double* out[5] = {p_out, rho_out, vx_out, vy_out, vz_out};
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 5; ++b)
            if ((const void*)in[a] == (const void*)out[b]) return HP_E_ARG;   // the sweep reads and writes apart
    unsigned char* base = nullptr;
    int rc = scratch_of_current_device(&base);
    if (rc) return rc;
    hp::Params* d_prm = (hp::Params*)(base + PRM_OFF);
    hp::DevStatus* d_st = (hp::DevStatus*)(base + ST_OFF);
    hipStream_t s = (hipStream_t)stream;

    hp::Params P;
    memcpy(&P, prm, sizeof(P));
    hipLaunchKernelGGL(scratch_init_kernel, dim3(1), dim3(64), 0, s, P, d_prm, d_st, newton_limit);
    if (hipGetLastError() != hipSuccess) return HP_E_HIP;

    hp::BoxView b;
    for (int f = 0; f < 5; ++f) {
        b.in[f] = in[f];
        b.out[f] = out[f];
    }
    b.gforce = gforce;
    b.prm = d_prm;
    b.st = d_st;
    b.nx = nx; b.ny = ny; b.nz = nz; b.nb = nbound;
    hp::SweepLaunch L;
    for (int f = 0; f < 4; ++f) L.old[f] = nullptr;
    L.fac = nullptr;
    L.dir = ndirection;
    L.ghost_mask = 7;
    L.do_speed = 0;
    L.seglen = 0;
    L.xcpw = 0;
    L.block = 256;
    L.precision = precision;
    L.grav = gforce != nullptr;
    L.cosmo = prm->dat != 0.0 || prm->at != 1.0 || prm->rat != 1.0;
    L.g53 = hp::params_g53(P);
    if (hp::launch_sweep(b, L, s) != hipSuccess) return HP_E_HIP;
    if (!counters) return HP_OK;

    hp::DevStatus st;
    if (hipMemcpyAsync(&st, d_st, sizeof(st), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return HP_E_HIP;
    memset(counters, 0, sizeof(*counters));
    counters->newton_iters_max = st.newton_max;
    counters->newton_nonconv = (int64_t)st.newton_nonconv;
    counters->neg_state_clamps = (int64_t)st.clamps;
    return st.newton_nonconv ? HP_E_NEWTON : HP_OK;
}
