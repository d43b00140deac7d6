// hp_nbody.hip -- the N-body step next to the hydro step in evolve_sys
// (hydropad/evolve_sys.F90:63-84): kick (nbody/kick.F90) with cell-in-cell
// interpolation of the static gravity array, drift (nbody/drift.F90), and
// the particle term of the CFL maximum (hydropad/speed.F90:43-56).
//
// Particles are the reference's ppos/pvel, (3,np) REAL(8) interleaved.  One
// thread per particle; a wave reads 64 consecutive particles, so the position
// and velocity loads and stores are contiguous 1.5 KB spans.  The kick's eight
// gforce corners are 4 pairs of 48 contiguous bytes.  The arithmetic keeps the
// reference's operation order (built without FP contraction), so results
// match the reference bit for bit.  evolve_sys runs kick, drift, kick over
// all particles with a static gforce; each particle's three updates depend
// on nothing else, so the step's form fuses them into one pass (one read and
// one write of the particle state instead of three).
#include <hip/hip_runtime.h>

#include "hp_core.h"
#include "hp_kernels.h"

namespace hp {

namespace {

struct Grid {
    const double* g;   // gforce (3,nx,ny,nz) or nullptr (no gravity: F = 0)
    int nx, ny, nz, nb;
    double fx, fy, fz; // real(ngridx..z) of the reference (FP32 -> FP64)
};

// nbody/kick.F90:47-117 on one rank: leading cell int(x*ngrid + 0.5) + nb
// (1-based), trilinear weights, 8-corner sum in the reference's order.
// False when the CIC cube leaves the array (the reference would read outside it).
__device__ __forceinline__ bool cic(const Grid& G, double x, double y, double z, double ff[3])
{
    const double x1g = (x - 0.0) * G.fx, x2g = (y - 0.0) * G.fy, x3g = (z - 0.0) * G.fz;
    const double t1 = x1g + 0.5, t2 = x2g + 0.5, t3 = x3g + 0.5;
    if (!(t1 > -1.0 && t2 > -1.0 && t3 > -1.0 && t1 < 1e9 && t2 < 1e9 && t3 < 1e9)) return false;
    const int i1 = (int)t1 + G.nb, j1 = (int)t2 + G.nb, k1 = (int)t3 + G.nb;
    if (i1 < 1 || j1 < 1 || k1 < 1 || i1 + 1 > G.nx || j1 + 1 > G.ny || k1 + 1 > G.nz) return false;
    if (!G.g) {
        ff[0] = ff[1] = ff[2] = 0.0;
        return true;
    }
    const double xi1 = (double)((float)(i1 - G.nb) - 0.5f);
    const double xi2 = (double)((float)(j1 - G.nb) - 0.5f);
    const double xi3 = (double)((float)(k1 - G.nb) - 0.5f);
    const double dd1 = x1g - xi1, dd2 = x2g - xi2, dd3 = x3g - xi3;
    const double de1 = 1.0 - dd1, de2 = 1.0 - dd2, de3 = 1.0 - dd3;
    const double d1 = de1 * de2 * de3, d2 = dd1 * de2 * de3, d3 = de1 * dd2 * de3, d4 = dd1 * dd2 * de3;
    const double d5 = de1 * de2 * dd3, d6 = dd1 * de2 * dd3, d7 = de1 * dd2 * dd3, d8 = dd1 * dd2 * dd3;
    const long sy = 3L * G.nx, sz = 3L * G.nx * G.ny;
    const double* c000 = G.g + 3L * (i1 - 1) + sy * (j1 - 1) + sz * (k1 - 1);
    const double* c010 = c000 + sy;
    const double* c001 = c000 + sz;
    const double* c011 = c001 + sy;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        ff[c] = c000[c] * d1 + c000[3 + c] * d2 + c010[c] * d3 + c010[3 + c] * d4 + c001[c] * d5 +
                c001[3 + c] * d6 + c011[c] * d7 + c011[3 + c] * d8;
    return true;
}

// kick.F90:122-124: v <- v (1 - dat dtat/2) + F dtat/2
__device__ __forceinline__ bool kick1(const Grid& G, double dat, double dtat, const double x[3], double v[3])
{
    double ff[3];
    if (!cic(G, x[0], x[1], x[2], ff)) return false;
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = v[c] * (1.0 - 0.5 * dat * dtat) + 0.5 * ff[c] * dtat;
    return true;
}

// drift.F90:24-32: x <- x + rdtath v, wrapped once into [0, boxsize)
__device__ __forceinline__ void drift1(double rdtath, const double bs[3], double x[3], const double v[3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] = x[c] + rdtath * v[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (x[c] < 0.0) x[c] = x[c] + bs[c];
        if (x[c] >= bs[c]) x[c] = x[c] - bs[c];
    }
}

__device__ __forceinline__ void flush_particles(DevStatus* st, unsigned long long outside, double vmax, bool speed)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        outside += __shfl_xor(outside, off, 64);
        if (speed) vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (outside) atomicAdd(&st->outside, outside);
        if (speed) atomicMax(&st->velmax_bits, (unsigned long long)__double_as_longlong(vmax));
    }
}

// what: 1 kick, 2 drift, 3 kick + drift + kick (the step's fused form).
// speed: also fold max |v| over every particle into the CFL maximum.
template <int WHAT>
__global__ void __launch_bounds__(256) particles_kernel(double* __restrict__ pos, double* __restrict__ vel, long np,
                                                        Grid G, const Params* __restrict__ prm, double dtat, double bs0,
                                                        double bs1, double bs2, DevStatus* st, int speed)
{
    const double dat = prm->dat, rdtath = prm->rdtath;
    const double bs[3] = {bs0, bs1, bs2};
    unsigned long long outside = 0;
    double vmax = 0.0;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < np; j += (long)gridDim.x * blockDim.x) {
        double x[3], v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { x[c] = pos[3 * j + c]; v[c] = vel[3 * j + c]; }
        if (x[0] != -1.0) {                               // kick.F90:37, drift.F90:23: inactive particle
            if ((WHAT & 1) && !kick1(G, dat, dtat, x, v)) ++outside;
            if (WHAT & 2) drift1(rdtath, bs, x, v);
            if (WHAT == 3 && !kick1(G, dat, dtat, x, v)) ++outside;
            if (WHAT & 2) {
#pragma unroll
                for (int c = 0; c < 3; ++c) pos[3 * j + c] = x[c];
            }
            if (WHAT & 1) {
#pragma unroll
                for (int c = 0; c < 3; ++c) vel[3 * j + c] = v[c];
            }
        }
        vmax = fmax(vmax, fmax(fmax(fabs(v[0]), fabs(v[1])), fabs(v[2])));   // speed.F90:43-47
    }
    flush_particles(st, outside, vmax, speed != 0);
}

__global__ void __launch_bounds__(256) particle_vmax_kernel(const double* __restrict__ vel, long np, DevStatus* st)
{
    double vmax = 0.0;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < np; j += (long)gridDim.x * blockDim.x)
        vmax = fmax(vmax, fmax(fmax(fabs(vel[3 * j]), fabs(vel[3 * j + 1])), fabs(vel[3 * j + 2])));
    flush_particles(st, 0, vmax, true);
}

long particle_grid(long np)
{
    long grid = (np + 255) / 256;
    if (grid > 65536) grid = 65536;
    return grid < 1 ? 1 : grid;
}

}  // namespace

double nbody_boxsize(int n, int nmax) { return (double)((float)n / (float)nmax); }

hipError_t launch_particles(int what, double* pos, double* vel, long np, const double* gforce, int nx, int ny, int nz,
                            int nb, const Params* prm, double dtat, DevStatus* st, int speed, hipStream_t s)
{
    if (np <= 0) return hipSuccess;
    Grid G;
    G.g = gforce;
    G.nx = nx; G.ny = ny; G.nz = nz; G.nb = nb;
    const int ngx = nx - 2 * nb, ngy = ny - 2 * nb, ngz = nz - 2 * nb;
    G.fx = (double)(float)ngx; G.fy = (double)(float)ngy; G.fz = (double)(float)ngz;
    const int nmax = max(ngx, max(ngy, ngz));
    const double b0 = nbody_boxsize(ngx, nmax), b1 = nbody_boxsize(ngy, nmax), b2 = nbody_boxsize(ngz, nmax);
    const dim3 grid((unsigned)particle_grid(np)), block(256);
    if (what == 1)
        hipLaunchKernelGGL(particles_kernel<1>, grid, block, 0, s, pos, vel, np, G, prm, dtat, b0, b1, b2, st, speed);
    else if (what == 2)
        hipLaunchKernelGGL(particles_kernel<2>, grid, block, 0, s, pos, vel, np, G, prm, dtat, b0, b1, b2, st, speed);
    else if (what == 3)
        hipLaunchKernelGGL(particles_kernel<3>, grid, block, 0, s, pos, vel, np, G, prm, dtat, b0, b1, b2, st, speed);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_particle_vmax(const double* vel, long np, DevStatus* st, hipStream_t s)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(particle_vmax_kernel, dim3((unsigned)particle_grid(np)), dim3(256), 0, s, vel, np, st);
    return hipGetLastError();
}

}  // namespace hp
