// hp_cosmo.h -- the Einstein-de Sitter step schedule of a cosmological box,
// shared by the host entry points and the device-resident step loop.
//   cosmo_init      hydropad/initialize_parameters.F90:136-178 (COSMO, omega_m = 1,
//                   not restarting from a dump)
//   cosmo_timedet   hydropad/functions.F90:38-151 (timedet_cosmo, omega_m = 1)
//   cosmo_factors   hydropad/expand.F90:18-20
// The reference's default-REAL literals are kept: 2.0/3.0 and -1.0/3.0 are
// FP32 quotients widened to double, 0.010 is the FP32 0.01.  Two departures,
// both because the reference's values cannot drive a run: the dt0 = 1e-9 cap
// (functions.F90:49,97) is not applied, and the particle speed (uninitialised
// vpart_max_pe, functions.F90:75-84) is taken as 0 -- there are no particles.
#pragma once
#include <math.h>

#include "hp_kernels.h"

namespace hp {

constexpr double K_TWO_THIRDS = (double)(2.0f / 3.0f);      // functions.F90:67-68
constexpr double K_MINUS_THIRD = (double)(-1.0f / 3.0f);    // functions.F90:68
constexpr double K_DT_INIT = (double)0.010f;                // initialize_parameters.F90:141

__host__ __device__ inline void cosmo_init(double a_start, double cour, double ca, double tstop, Cosmo& s)
{
    s.cour = cour;
    s.ca = ca;
    s.tstop = tstop;
    const double tin = pow(a_start, 1.5);
    const double told = tin;
    s.dt = K_DT_INIT * tin;
    s.t = tin + s.dt;
    const double thalf = told + s.dt / 2.0;
    s.at = pow(told, K_TWO_THIRDS);
    s.atnew = pow(s.t, K_TWO_THIRDS);
    s.dat = K_TWO_THIRDS * pow(told, K_MINUS_THIRD);
    s.ath = pow(told + s.dt / 2., K_TWO_THIRDS);
    s.dath = K_TWO_THIRDS * pow(thalf, K_MINUS_THIRD);
    s.datnew = K_TWO_THIRDS * pow(s.t, K_MINUS_THIRD);
    s.rat = 1.0 / s.at;
    s.rdtath = s.dt / s.ath;
    s.nstep = 0;
}

// Called after step s.nstep+1 with that step's CFL maximum: marks it done
// and sets up the next one.
__host__ __device__ inline void cosmo_timedet(Cosmo& s, double velmax, double dx)
{
    s.nstep += 1;
    const double told = s.t;
    const double rdt = (s.nstep > 1) ? 5.0 / (6.0 * s.dt) : 0.0;
    s.at = pow(s.t, K_TWO_THIRDS);
    s.dat = K_TWO_THIRDS * pow(s.t, K_MINUS_THIRD);
    const double dta = pow((1.0 + s.ca) * s.at, 1.5) - s.t;
    s.rat = 1.0 / s.at;
    const double rdtc = velmax / (s.cour * dx * s.at);
    double dt = 1.0 / (rdt > rdtc ? rdt : rdtc);
    dt = dt < dta ? dt : dta;
    if (dt >= 10.0) dt = 0.1;
    double t = s.t + dt;
    if (t > s.tstop) {
        t = s.tstop;
        dt = t - told;
    }
    s.t = t;
    s.dt = dt;
    const double dth = 0.5 * dt;
    const double thalf = told + dth;
    s.atnew = pow(t, K_TWO_THIRDS);
    s.ath = pow(thalf, K_TWO_THIRDS);
    s.dath = K_TWO_THIRDS * pow(thalf, K_MINUS_THIRD);
    s.datnew = K_TWO_THIRDS * pow(t, K_MINUS_THIRD);
    s.rdtath = dt / s.ath;
}

__host__ __device__ inline void cosmo_factors(const Cosmo& s, double fac[3])
{
    const double afact = s.dt * (s.dath / s.ath);
    fac[0] = afact;
    fac[1] = 1.0 / (1.0 + afact);
    fac[2] = 1.0 / (2.0 + afact);
}

}  // namespace hp
