// hp_kernels.hip -- the small kernels of the step (periodic ghost fill, CFL
// maximum, timedet) and the dispatcher over the sweep variants of
// hp_kernels_exact.hip / hp_kernels_fast.hip.
#include "hp_core.h"
#include "hp_cosmo.h"
#include "hp_kernels.h"

namespace hp {

hipError_t launch_sweep_exact(const BoxView& b, const SweepLaunch& L, hipStream_t s);
hipError_t launch_sweep_fast(const BoxView& b, const SweepLaunch& L, hipStream_t s);

// ============================================================================
// Periodic ghost fill of one axis for the 5 fields, transverse range complete
// (ghosts included), as the non-MPI branch of hydropad/exchange_mesh.F90:47-123.
__global__ void __launch_bounds__(256) fill_axis(double* f0, double* f1, double* f2, double* f3, double* f4, int nx,
                                                 int ny, int nz, int nb, int dir)
{
    long n_t1, n_t2, s_ax, s_t1, s_t2;
    int n_ax;
    if (dir == 1) { n_ax = nx; n_t1 = ny; n_t2 = nz; s_ax = 1; s_t1 = nx; s_t2 = (long)nx * ny; }
    else if (dir == 2) { n_ax = ny; n_t1 = nx; n_t2 = nz; s_ax = nx; s_t1 = 1; s_t2 = (long)nx * ny; }
    else { n_ax = nz; n_t1 = nx; n_t2 = ny; s_ax = (long)nx * ny; s_t1 = 1; s_t2 = nx; }
    const long tot = 2L * nb * n_t1 * n_t2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tot) return;
    const long a = t % n_t1;
    const long r = t / n_t1;
    const long b = r % n_t2;
    const int g = (int)(r / n_t2);                 // 0..2nb-1
    const int N = n_ax - 2 * nb;
    int dst, src;
    if (g < nb) { dst = g; src = g + N; }          // low ghosts <- high interior
    else { dst = N + g; src = g; }                 // high ghosts (N+nb..) <- low interior (nb..)
    const long off = a * s_t1 + b * s_t2;
    const long qd = off + dst * s_ax, qs = off + src * s_ax;
    f0[qd] = f0[qs]; f1[qd] = f1[qs]; f2[qd] = f2[qs]; f3[qd] = f3[qs]; f4[qd] = f4[qs];
}

// CFL maximum over the interior, hydropad/speed.F90:22-31.
__global__ void __launch_bounds__(256) speed_kernel(const double* __restrict__ p, const double* __restrict__ rho,
                                                    const double* __restrict__ vx, const double* __restrict__ vy,
                                                    const double* __restrict__ vz, int nx, int ny, int nz, int nb,
                                                    const Params* __restrict__ prm, DevStatus* st)
{
    const int Nx = nx - 2 * nb, Ny = ny - 2 * nb, Nz = nz - 2 * nb;
    const long tot = (long)Nx * Ny * Nz;
    const double gamma = prm->gamma;
    double vmax = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long i = t % Nx, r = t / Nx;
        const long j = r % Ny, k = r / Ny;
        const long q = (nb + i) + (long)nx * ((nb + j) + (long)ny * (nb + k));
        const double vs = sqrt(gamma * p[q] / rho[q]);
        vmax = tmax(vmax, tmax(tmax(fabs(vx[q]), fabs(vy[q])), fabs(vz[q])) + vs);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&st->velmax_bits, (unsigned long long)__double_as_longlong(vmax));
}

// timedet, hydropad/functions.F90:2-35 (non-cosmological branch).
__global__ void timedet_kernel(Params* prm, DevStatus* st, double cour, double tstop)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double velmax = __longlong_as_double((long long)st->velmax_bits);
    const double told = st->t;
    const double rdtc = velmax / (cour * prm->dx);
    double dt = (velmax == 0.0) ? 0.1 : 1.0 / rdtc;
    double t = told + dt;
    if (t > tstop) { t = tstop; dt = t - told; }
    st->t = t;
    prm->at = 1.0;
    prm->dat = 0.0;
    prm->rat = 1.0;
    prm->dt = dt;
    prm->rdtath = dt / 1.0;
    st->velmax_bits = 0ull;
}

// timedet_cosmo on the device; writes the scalars of the next step.
__global__ void timedet_cosmo_kernel(Params* prm, DevStatus* st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double velmax = __longlong_as_double((long long)st->velmax_bits);
    Cosmo c = st->cosmo;
    cosmo_timedet(c, velmax, prm->dx);
    st->cosmo = c;
    prm->at = c.at;
    prm->dat = c.dat;
    prm->rat = c.rat;
    prm->dt = c.dt;
    prm->rdtath = c.rdtath;
    cosmo_factors(c, st->fac);
    st->velmax_bits = 0ull;
}

// expand.F90:22-38 over the whole array (the reference loops over ghosts too).
__global__ void __launch_bounds__(256) expand_kernel(double* __restrict__ p, double* __restrict__ vx,
                                                     double* __restrict__ vy, double* __restrict__ vz,
                                                     const double* __restrict__ po, const double* __restrict__ vxo,
                                                     const double* __restrict__ vyo, const double* __restrict__ vzo,
                                                     long n, double afact, double afact1, double afact2)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        vx[i] = (2.0 * vx[i] - afact * vxo[i]) * afact2;
        vy[i] = (2.0 * vy[i] - afact * vyo[i]) * afact2;
        vz[i] = (2.0 * vz[i] - afact * vzo[i]) * afact2;
        const double ppp = p[i];
        double pn = (p[i] - afact * po[i]) * afact1;
        if (pn <= 0.0) pn = (ppp > 0.0) ? ppp : po[i];
        p[i] = pn;
    }
}

struct CopyArgs {
    double* dst[5];
    const double* src[5];
};

__global__ void __launch_bounds__(256) copy_kernel(CopyArgs a, int nfield, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        for (int f = 0; f < nfield; ++f) a.dst[f][i] = a.src[f][i];
}

// what: bit 0 clears the CFL maximum, bit 1 the counters.
__global__ void zero_status_kernel(DevStatus* st, int what)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (what & 1) st->velmax_bits = 0ull;
        if (what & 2) {
            st->newton_max = 0;
            st->newton_nonconv = 0ull;
            st->clamps = 0ull;
            st->outside = 0ull;
        }
    }
}

// ============================================================================
// host launchers

hipError_t launch_sweep(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    if (L.dir < 1 || L.dir > 3) return hipErrorInvalidValue;
    if (L.precision == PREC_FP64_EXACT) return launch_sweep_exact(b, L, s);
    return launch_sweep_fast(b, L, s);
}

hipError_t launch_fill(double* const f[5], int nx, int ny, int nz, int nb, int dir, hipStream_t s)
{
    const int order[3] = {3, 1, 2};   // exchange_mesh.F90:47-123: z, x, y
    for (int o = 0; o < 3; ++o) {
        const int d = order[o];
        if (dir != 0 && dir != d) continue;
        long nt1, nt2;
        if (d == 1) { nt1 = ny; nt2 = nz; }
        else if (d == 2) { nt1 = nx; nt2 = nz; }
        else { nt1 = nx; nt2 = ny; }
        const long tot = 2L * nb * nt1 * nt2;
        const long grid = (tot + 255) / 256;
        hipLaunchKernelGGL(fill_axis, dim3((unsigned)grid), dim3(256), 0, s, f[0], f[1], f[2], f[3], f[4], nx, ny, nz,
                           nb, d);
    }
    return hipGetLastError();
}

hipError_t launch_speed(const double* const f[5], int nx, int ny, int nz, int nb, const Params* prm, DevStatus* st,
                        hipStream_t s)
{
    const long tot = (long)(nx - 2 * nb) * (ny - 2 * nb) * (nz - 2 * nb);
    long grid = (tot + 255) / 256;
    if (grid > 8192) grid = 8192;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(speed_kernel, dim3((unsigned)grid), dim3(256), 0, s, f[0], f[1], f[2], f[3], f[4], nx, ny, nz,
                       nb, prm, st);
    return hipGetLastError();
}

hipError_t launch_timedet(Params* prm, DevStatus* st, double cour, double tstop, hipStream_t s)
{
    hipLaunchKernelGGL(timedet_kernel, dim3(1), dim3(64), 0, s, prm, st, cour, tstop);
    return hipGetLastError();
}

hipError_t launch_timedet_cosmo(Params* prm, DevStatus* st, hipStream_t s)
{
    hipLaunchKernelGGL(timedet_cosmo_kernel, dim3(1), dim3(64), 0, s, prm, st);
    return hipGetLastError();
}

static long stream_grid(long n)
{
    long grid = (n + 255) / 256;
    if (grid > 16384) grid = 16384;
    return grid < 1 ? 1 : grid;
}

hipError_t launch_expand(double* const f[5], const double* const old[4], long ncell, double afact, hipStream_t s)
{
    const double afact1 = 1.0 / (1.0 + afact), afact2 = 1.0 / (2.0 + afact);
    hipLaunchKernelGGL(expand_kernel, dim3((unsigned)stream_grid(ncell)), dim3(256), 0, s, f[0], f[2], f[3], f[4],
                       old[0], old[1], old[2], old[3], ncell, afact, afact1, afact2);
    return hipGetLastError();
}

hipError_t launch_copy(double* const* dst, const double* const* src, int nfield, long ncell, hipStream_t s)
{
    if (nfield < 1 || nfield > 5) return hipErrorInvalidValue;
    CopyArgs a;
    for (int f = 0; f < 5; ++f) {
        a.dst[f] = f < nfield ? dst[f] : nullptr;
        a.src[f] = f < nfield ? src[f] : nullptr;
    }
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)stream_grid(ncell)), dim3(256), 0, s, a, nfield, ncell);
    return hipGetLastError();
}

hipError_t launch_zero_status(DevStatus* st, int what, hipStream_t s)
{
    hipLaunchKernelGGL(zero_status_kernel, dim3(1), dim3(64), 0, s, st, what);
    return hipGetLastError();
}

}  // namespace hp
