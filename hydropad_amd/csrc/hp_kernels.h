// hp_kernels.h -- host-side launch interface of the sweep, ghost-fill,
// speed and timedet kernels (hp_kernels.hip).  Host code only includes this.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hp {

struct Params;

// Einstein-de Sitter time-step schedule of a cosmological box
// (hydropad/functions.F90:38-151, hydropad/initialize_parameters.F90:136-178);
// same layout as hp_cosmo in include/hydropad_gpu.h.
struct Cosmo {
    double cour, ca, tstop;
    double t, dt;
    double at, dat, ath, dath, atnew, datnew;
    double rat, rdtath;
    long long nstep;
};

// Device-resident per-box step state and counters.
struct DevStatus {
    unsigned long long velmax_bits;   // max(|v|+cs) as the bit pattern of a non-negative double
    int newton_max;
    int newton_cap;                    // Newton iteration limit; 0 = the reference's nnewtmax = 100 (riemann.F90:51)
    unsigned long long newton_nonconv;
    unsigned long long clamps;
    unsigned long long outside;        // particle kicks skipped: CIC cube outside the array
    double t;                          // time at the end of the upcoming step (timedet semantics)
    double fac[3];                     // expansion factors of the upcoming step (expand.F90:18-20)
    Cosmo cosmo;                       // schedule of a cosmological box
    // run control of a box in a batch (unused by a single context)
    long long nstep_done;              // steps completed since the box was added
    long long nstep_target;            // steps to run
    int done;                          // nstep_done reached nstep_target
    int comoving;
    double cour, tstop;                // non-comoving timedet inputs
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

// True when the scalars are gamma = 5/3 with exactly the derived constants
// hp_params_init computes (then the constant-folded kernels give identical results).
bool params_g53(const Params& P);

struct SweepLaunch {
    int dir;          // 1,2,3
    int ghost_mask;   // bit d-1: also write the periodic face-ghost images along axis d
    int do_speed;     // fuse the CFL maximum into this sweep
    int seglen;       // streaming segment length for y/z sweeps (<= 0: whole pencil)
    int xcpw;         // x sweep: chunks per wave (<= 0: about 16 resident rounds)
    int block;        // threads per block
    int precision;    // Precision
    bool grav, cosmo;
    bool g53;         // gamma == 5/3 with the derived scalars of hp_params_init: constant-folded kernels
    // expansion step fused into this sweep (hydropad/expand.F90:18-41):
    // start-of-step p, vx, vy, vz, or old[0] == nullptr; fac -> device
    // (afact, 1/(1+afact), 1/(2+afact))
    const double* old[4];
    const double* fac;
};

hipError_t launch_sweep(const BoxView& b, const SweepLaunch& L, hipStream_t s);
// Periodic ghost fill of the 5 fields of `f` (in place); dir 1..3 or 0 = all (z, x, y).
hipError_t launch_fill(double* const f[5], int nx, int ny, int nz, int nb, int dir, hipStream_t s);
hipError_t launch_speed(const double* const f[5], int nx, int ny, int nz, int nb, const Params* prm,
                        DevStatus* st, hipStream_t s);
// timedet (hydropad/functions.F90:2-35) on the device: consumes velmax_bits,
// writes prm->dt, prm->rdtath, st->t, and zeroes velmax_bits.
hipError_t launch_timedet(Params* prm, DevStatus* st, double cour, double tstop, hipStream_t s);
// timedet_cosmo (EdS branch, hydropad/functions.F90:38-151) on the device:
// consumes velmax_bits, advances st->cosmo, writes prm->{at,dat,rat,dt,rdtath}
// and st->fac for the next step's expansion, zeroes velmax_bits.
hipError_t launch_timedet_cosmo(Params* prm, DevStatus* st, hipStream_t s);
// expand (hydropad/expand.F90:18-41) over the whole array, in place, against
// the start-of-step p, vx, vy, vz in old[].
hipError_t launch_expand(double* const f[5], const double* const old[4], long ncell, double afact, hipStream_t s);
// dst[k][i] = src[k][i] for the n fields listed (a plain streaming copy).
hipError_t launch_copy(double* const* dst, const double* const* src, int nfield, long ncell, hipStream_t s);
// what: bit 0 clears the CFL maximum, bit 1 the counters.
hipError_t launch_zero_status(DevStatus* st, int what, hipStream_t s);

// N-body (hp_nbody.hip): what = 1 kick (nbody/kick.F90), 2 drift
// (nbody/drift.F90), 3 kick + drift + kick (one step of
// hydropad/evolve_sys.F90:63-84).  pos/vel are (3,np); gforce may be nullptr
// (zero force).  speed != 0 folds max |v| into st->velmax_bits
// (hydropad/speed.F90:43-56); skipped kicks count in st->outside.
hipError_t launch_particles(int what, double* pos, double* vel, long np, const double* gforce, int nx, int ny, int nz,
                            int nb, const Params* prm, double dtat, DevStatus* st, int speed, hipStream_t s);
hipError_t launch_particle_vmax(const double* vel, long np, DevStatus* st, hipStream_t s);

}  // namespace hp
