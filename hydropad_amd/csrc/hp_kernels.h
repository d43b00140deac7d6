// hp_kernels.h -- host-side launch interface of the sweep, ghost-fill,
// speed and timedet kernels (hp_kernels.hip).  Host code only includes this.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hp {

struct Params;

// Device-resident per-box step state and counters.
struct DevStatus {
    unsigned long long velmax_bits;   // max(|v|+cs) as the bit pattern of a non-negative double
    int newton_max;
    int pad0;
    unsigned long long newton_nonconv;
    unsigned long long clamps;
    double t;                          // time at the end of the upcoming step (timedet semantics)
    double pad1[3];
};

// One box as seen by a sweep.  Fields are (nx,ny,nz) arrays with pitch nx,
// cell (0,0,0) the first ghost cell; interior extents nx-2nb etc.
struct BoxView {
    const double* in[5];    // p, rho, vx, vy, vz
    double* out[5];
    const double* gforce;   // (3,nx,ny,nz) interleaved, or nullptr
    Params* prm;            // device parameter block of this box
    DevStatus* st;
    int nx, ny, nz, nb;
};

enum Precision { PREC_FP64 = 0, PREC_MIXED = 1, PREC_FP64_EXACT = 2 };

struct SweepLaunch {
    int dir;          // 1,2,3
    int ghost_mask;   // bit d-1: also write the periodic face-ghost images along axis d
    int do_speed;     // fuse the CFL maximum into this sweep
    int seglen;       // streaming segment length for y/z sweeps (<= 0: whole pencil)
    int block;        // threads per block
    int precision;    // Precision
    bool grav, cosmo;
};

hipError_t launch_sweep(const BoxView& b, const SweepLaunch& L, hipStream_t s);
// Periodic ghost fill of the 5 fields of `f` (in place); dir 1..3 or 0 = all (z, x, y).
hipError_t launch_fill(double* const f[5], int nx, int ny, int nz, int nb, int dir, hipStream_t s);
hipError_t launch_speed(const double* const f[5], int nx, int ny, int nz, int nb, const Params* prm,
                        DevStatus* st, hipStream_t s);
// timedet (hydropad/functions.F90:2-35) on the device: consumes velmax_bits,
// writes prm->dt, prm->rdtath, st->t, and zeroes velmax_bits.
hipError_t launch_timedet(Params* prm, DevStatus* st, double cour, double tstop, hipStream_t s);
// what: bit 0 clears the CFL maximum, bit 1 the counters.
hipError_t launch_zero_status(DevStatus* st, int what, hipStream_t s);

}  // namespace hp
