// hp_core.h -- device math of one directionally split PPM sweep, cell-centric.
//
// The reference (hydro/order2interp.F90, hydro/ppm_functs.F90, hydro/dda.F90)
// evaluates, for every interface, the parabolas of BOTH neighbouring cells
// (its loop i=nmin..nmax runs twice, order2interp.F90:42).  Every quantity it
// builds for the left state of interface i-1/2 depends only on cell i-1 and
// its 5-cell window, and every quantity of the right state only on cell i:
//   parabola(c)   <- a(c-2..c+2), flat(c) from the pairs in c-2..c+2
//   right face of c (the reference's "left" state of interface c+1/2)
//   left  face of c (the reference's "right" state of interface c-1/2)
// so here each cell's parabola and both traced face states are computed once,
// and each interface only runs the Riemann solver.  The arithmetic of each
// function keeps the reference's operation order; built with
// -ffp-contract=off (HP_EXACT) it reproduces the reference bit for bit except
// where OCML's pow differs from glibc's in the last bit.
//
// Real type R is double for the FP64 form and float for the reconstruction of
// the mixed form; the Newton iteration and the dual-energy update are always
// double (see hp_riemann_* and hp_integrate).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hp {

// Device mirror of hp_params (include/hydropad_gpu.h); identical layout.
struct Params {
    double gamma, gamma1, rgamma, rgamma1, m, rm, gf;
    double dt, dx, rdx, rdtath, at, dat, rat;
    double flatvalue, mingradflat, eta1, eta2, dmax;
};

// The reference's default-REAL literals inside REAL(8) expressions, widened.
constexpr double K_SIXTH = (double)0.16666667f;  // ppm_functs.F90:33-34, dda.F90:198-199
constexpr double K_THIRD = (double)0.33333333f;  // ppm_functs.F90:103,122
constexpr double K_0010 = (double)0.010f;        // dda.F90:200
constexpr double K_0050 = (double)0.050f;        // dda.F90:203
constexpr double K_1EM15 = (double)1e-15f;       // order2interp.F90:176-189
constexpr double K_TERR = (double)1e-3f;         // riemann.F90:48
constexpr int K_NNEWTMAX = 100;                   // riemann.F90:51

template <class T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }

// Variable slots along a sweep: pressure, density, normal velocity, the two
// transverse velocities, and (with gravity) the gravity component.
enum { VP = 0, VR = 1, VU = 2, VV = 3, VW = 4, VG = 5 };

// Scalars of one sweep, loaded once per kernel from the device parameter
// block and converted to the reconstruction precision R.
template <class R> struct SweepScalars {
    R gamma, dt, rat, dx, rdx, dat, at, flatvalue, mingradflat;
    R dtrat;   // dt*rat, used only by the fast forms
    R cdat2;   // 2*dat*rat*dt (expansion term of beta/beta0), used only by the fast forms
};

// ---------------------------------------------------------------------------
// Reconstruction (hydro/ppm_functs.F90, hydro/dda.F90)

// MC-limited slope, hydro/ppm_functs.F90:167-181
template <class R> __device__ __forceinline__ R ajside(R am1, R a, R ap1)
{
    if ((ap1 - a) * (a - am1) > R(0)) {
        R da = R(0.5) * (ap1 - am1);
        R mn = tmin(tmin(fabs(da), R(2) * fabs(a - am1)), R(2) * fabs(ap1 - a));
        return mn * da / fabs(da);
    }
    return R(0);
}

// Interface value between cells (c-1 | c) before limiting: the reference's
// ajl(2) of cell c, identical to ajr(2) of cell c-1 (ppm_functs.F90:33-34).
template <class R> __device__ __forceinline__ R iface_value(R am1, R a0, R dm1, R d0)
{
    return am1 + R(0.5) * (a0 - am1) + (dm1 - d0) * R(K_SIXTH);
}

// One pair of the flattening test (order2interp.F90:65-113): true when
// |a1-a0|/min(a1,a0) >= thr.  A pair with a1*a0 == 0 contributes nothing.
template <class R> __device__ __forceinline__ bool pair_steep(R a0, R a1, R thr)
{
    if (a1 * a0 != R(0)) {
        R gl = fabs((a1 - a0) / tmin(a1, a0));
        return gl >= thr;
    }
    return false;
}

// Contact steepening of the density parabola, hydro/dda.F90:182-216 and its
// use in ppm_functs.F90:35-40.  a1..a5 = rho(c-2..c+2); dmalf/dmarg the
// slopes of cells c-1 and c+1.
template <class R>
__device__ __forceinline__ void steepen(R a1, R a2, R a3, R a4, R a5, R dmalf, R dmarg, R dx,
                                        R& ajl, R& ajr)
{
    R deltam2 = R(K_SIXTH) * (a3 - R(2) * a2 + a1) / (dx * dx);
    R deltap2 = R(K_SIXTH) * (a5 - R(2) * a4 + a3) / (dx * dx);
    R agrad = fabs(a4 - a2) - R(K_0010) * tmin(fabs(a4), fabs(a2));
    R eta;
    if (-deltam2 * deltap2 > R(0) && agrad > R(0)) {
        R e = -(deltap2 - deltam2) * dx * dx / (a4 - a2);
        eta = tmax(R(0), tmin(R(20) * (e - R(K_0050)), R(1)));
    } else {
        eta = R(0);
    }
    R ajld = a2 + R(0.5) * dmalf;
    R ajrd = a4 - R(0.5) * dmarg;
    ajl = (R(1) - eta) * ajl + eta * ajld;
    ajr = (R(1) - eta) * ajr + eta * ajrd;
}

template <class R> struct Parab { R l, r, d, a6; };

// Monotonicity + flattening, ppm_functs.F90:41-54
template <class R> __device__ __forceinline__ Parab<R> limit_parab(R a3, R ajl, R ajr, R flat)
{
    R d = ajr - ajl;
    R a6 = R(6) * (a3 - R(0.5) * (ajl + ajr));
    if ((ajr - a3) * (a3 - ajl) <= R(0)) {
        ajl = a3;
        ajr = a3;
    } else if (d * a6 > d * d) {
        ajl = R(3) * a3 - R(2) * ajr;
    } else if (d * a6 < (-(d * d))) {
        ajr = R(3) * a3 - R(2) * ajl;
    }
    ajl = flat * a3 + (R(1) - flat) * ajl;
    ajr = flat * a3 + (R(1) - flat) * ajr;
    Parab<R> q;
    q.l = ajl;
    q.r = ajr;
    q.d = ajr - ajl;
    q.a6 = R(6) * (a3 - R(0.5) * (ajl + ajr));
    return q;
}

// Averages of the parabola over a distance x*dx from its right edge (fl) or
// left edge (fr), ppm_functs.F90:91-124; x = y/dx is computed by the caller.
template <class R> __device__ __forceinline__ R pfl(const Parab<R>& q, R x)
{
    return q.r - R(0.5) * x * (q.d - (R(1) - R(2) * x * R(K_THIRD)) * q.a6);
}
template <class R> __device__ __forceinline__ R pfr(const Parab<R>& q, R x)
{
    return q.l + R(0.5) * x * (q.d + (R(1) - R(2) * x * R(K_THIRD)) * q.a6);
}

// ---------------------------------------------------------------------------
// Characteristic tracing of one face (order2interp.F90:174-258)

template <class R> struct Face { R p, r, u, v, w; };

// beta, ppm_functs.F90:129-144.  COSMO=false drops the dat-terms, which are
// exact zeros when dat == 0; GRAV=false drops gc (g == 0).
template <class R, bool GRAV, bool COSMO>
__device__ __forceinline__ R beta_fn(const SweepScalars<R>& S, R ct, R ut, R uc, R pt, R pc, R gc, R sign)
{
    R sum = (ut - uc) + sign * (pt - pc) / ct;
    if (COSMO) sum = sum + sign * S.dt * (R(2) * S.dat * pc / (ct * S.at));
    if (GRAV || COSMO) {
        R gvc = ((GRAV ? -gc : R(0)) + (COSMO ? S.dat * uc : R(0))) * S.rat;
        sum = sum + S.dt * gvc;
    }
    return -sign * sum / (R(2) * ct);
}

// beta0, ppm_functs.F90:149-162
template <class R, bool COSMO>
__device__ __forceinline__ R beta0_fn(const SweepScalars<R>& S, R ct, R rhot, R rhoc, R pt, R pc)
{
    R b = (pt - pc) / (ct * ct) + R(1) / rhot - R(1) / rhoc;
    if (COSMO) b = b + (R(2) * S.dat * pc) / (S.at * ct * ct) * S.dt;
    return b;
}

// Traced state at the RIGHT face of cell c (the reference's "left" state of
// interface c+1/2): fl of c's parabolas at the +,0,- characteristics with the
// cell's normal velocity u and sound speed cs.  q[] holds the parabolas of
// p, rho, u, v, w (and g).  nclamp counts negative-state clamps.
template <class R, bool GRAV, bool COSMO>
__device__ __forceinline__ Face<R> trace_right_face(const SweepScalars<R>& S, const Parab<R>* q, R u,
                                                    R cs, int& nclamp)
{
    const R dt = S.dt, rat = S.rat, dx = S.dx;
    R y3 = dt * (u + cs) * rat;                       // ppm_functs.F90:59,74
    R xt = tmax(y3, R(0)) / dx;
    R x1 = (dt * (u - cs) * rat) / dx;                // :70
    R x2 = (dt * u * rat) / dx;                       // :72
    R x3 = y3 / dx;
    R ptl = pfl(q[VP], xt), rtl = pfl(q[VR], xt), utl = pfl(q[VU], xt);
    R vtl = pfl(q[VV], xt), wtl = pfl(q[VW], xt);
    R pc1 = pfl(q[VP], x1), pc2 = pfl(q[VP], x2), pc3 = pfl(q[VP], x3);
    R rc2 = pfl(q[VR], x2), uc1 = pfl(q[VU], x1), uc3 = pfl(q[VU], x3);
    R vc2 = pfl(q[VV], x2), wc2 = pfl(q[VW], x2);
    R gc1 = R(0), gc3 = R(0);
    if (GRAV) { gc1 = pfl(q[VG], x1); gc3 = pfl(q[VG], x3); }
    if (ptl < R(0)) { ptl = R(K_1EM15); ++nclamp; }
    if (rtl < R(0)) { rtl = R(K_1EM15); ++nclamp; }
    R ct = sqrt(S.gamma * ptl * rtl);
    R lm = (u - cs) * rat, l0 = u * rat, lp = (u + cs) * rat;
    R bm = (lm <= R(0)) ? R(0) : beta_fn<R, GRAV, COSMO>(S, ct, utl, uc1, ptl, pc1, gc1, R(-1));
    R bp = (lp <= R(0)) ? R(0) : beta_fn<R, GRAV, COSMO>(S, ct, utl, uc3, ptl, pc3, gc3, R(1));
    R b0 = R(0);
    R vf = vtl, wf = wtl;
    if (l0 > R(0)) {
        b0 = beta0_fn<R, COSMO>(S, ct, rtl, rc2, ptl, pc2);
        vf = vc2;
        wf = wc2;
    }
    Face<R> f;
    f.p = ptl + ct * ct * (bp + bm);
    if (f.p <= R(0)) f.p = ptl;
    f.u = utl + ct * (bp - bm);
    f.r = R(1) / (R(1) / rtl - bm - b0 - bp);
    if (f.r <= R(0)) f.r = rtl;
    f.v = vf;
    f.w = wf;
    return f;
}

// Traced state at the LEFT face of cell c (the reference's "right" state of
// interface c-1/2).
template <class R, bool GRAV, bool COSMO>
__device__ __forceinline__ Face<R> trace_left_face(const SweepScalars<R>& S, const Parab<R>* q, R u,
                                                   R cs, int& nclamp)
{
    const R dt = S.dt, rat = S.rat, dx = S.dx;
    R y1 = -dt * (u - cs) * rat;                      // ppm_functs.F90:62,79
    R xt = tmax(y1, R(0)) / dx;
    R x1 = y1 / dx;
    R x2 = (-dt * u * rat) / dx;                      // :81
    R x3 = (-dt * (u + cs) * rat) / dx;               // :83
    R ptr = pfr(q[VP], xt), rtr = pfr(q[VR], xt), utr = pfr(q[VU], xt);
    R vtr = pfr(q[VV], xt), wtr = pfr(q[VW], xt);
    R pc1 = pfr(q[VP], x1), pc2 = pfr(q[VP], x2), pc3 = pfr(q[VP], x3);
    R rc2 = pfr(q[VR], x2), uc1 = pfr(q[VU], x1), uc3 = pfr(q[VU], x3);
    R vc2 = pfr(q[VV], x2), wc2 = pfr(q[VW], x2);
    R gc1 = R(0), gc3 = R(0);
    if (GRAV) { gc1 = pfr(q[VG], x1); gc3 = pfr(q[VG], x3); }
    if (ptr < R(0)) { ptr = R(K_1EM15); ++nclamp; }
    if (rtr < R(0)) { rtr = R(K_1EM15); ++nclamp; }
    R ct = sqrt(S.gamma * ptr * rtr);
    R lm = (u - cs) * rat, l0 = u * rat, lp = (u + cs) * rat;
    R bp = (lp >= R(0)) ? R(0) : beta_fn<R, GRAV, COSMO>(S, ct, utr, uc3, ptr, pc3, gc3, R(1));
    R bm = (lm >= R(0)) ? R(0) : beta_fn<R, GRAV, COSMO>(S, ct, utr, uc1, ptr, pc1, gc1, R(-1));
    R b0 = R(0);
    R vf = vtr, wf = wtr;
    if (l0 < R(0)) {
        b0 = beta0_fn<R, COSMO>(S, ct, rtr, rc2, ptr, pc2);
        vf = vc2;
        wf = wc2;
    }
    Face<R> f;
    f.p = ptr + ct * ct * (bp + bm);
    if (f.p <= R(0)) f.p = ptr;
    f.u = utr + ct * (bp - bm);
    f.r = R(1) / (R(1) / rtr - bm - b0 - bp);
    if (f.r <= R(0)) f.r = rtr;
    f.v = vf;
    f.w = wf;
    return f;
}

// ---------------------------------------------------------------------------
// Riemann solver (hydro/riemann.F90, hydro/riemannsolver.F90), always FP64.

struct RScal {
    double gamma, gamma1, rgamma, rgamma1, m, rm, gf, dadt, dmax;
    int m5;   // m == 1/5 (gamma = 5/3): powers of m, 1/gamma, 1/m are roots/products
    int nmax; // iterations after which the reference stops (nnewtmax = 100, riemann.F90:51)
    int nskip; // K_NNEWTMAX - nmax: the fast iteration's head start for a limit other than 100
};

__device__ __forceinline__ double rarwave(double f1, double f2, double vt, double vh)
{
    return f2 + vt * (f1 - f2) / (vt - vh);               // riemann.F90:7
}

__device__ __forceinline__ double rhoshock(const RScal& R, double p2, double p1, double rho1)
{
    return rho1 * (R.gamma1 * p2 / p1 + 1.0) / (R.gamma1 + p2 / p1);   // riemann.F90:22
}

// van Leer Newton iteration, riemann.F90:29-111.  Returns iterations or -1
// where the reference stops (more than 100).
__device__ __forceinline__ int newton(const RScal& R, double& pstar, double& vstar, double pleft,
                                      double pright, double rholeft, double rhoright, double vxleft,
                                      double vxright)
{
    const double gamma = R.gamma, m = R.m, gf = R.gf;
    double err = 1.e10, a = 0.0, b = 0.0;
    int nnewt = 0;
    double kappa = sqrt(gamma * (pleft + pright) * (rholeft + rhoright) * 0.25);
    double pcheck = 0.50 * (pleft + pright) + 0.50 * kappa * (vxleft + vxright);
    if (pcheck > 0) pstar = pcheck;
    const double cleft = sqrt(gamma * pleft * rholeft);
    const double cright = sqrt(gamma * pright * rhoright);
    while (err > K_TERR) {
        nnewt = nnewt + 1;
        if (nnewt > R.nmax) return -1;
        double pstarold = pstar;
        double pstup = (1. - pow(pstar / pleft, m));
        if (pstar >= pleft) a = cleft * sqrt(gf * pstar / pleft + m);
        else if (pstup == 0.) a = cleft * sqrt(gf * pstar / pleft + m);
        else a = m * cleft * (1.0 - pstar / pleft) / (1 - pow(pstar / pleft, m));
        pstup = (1. - pow(pstar / pright, m));
        if (pstar >= pright) b = cright * sqrt(gf * pstar / pright + m);
        else if (pstup == 0.) b = cright * sqrt(gf * pstar / pright + m);
        else b = m * cright * (1.0 - pstar / pright) / (1 - pow(pstar / pright, m));
        double vprimel, vprimer;
        if (pstar >= pleft) vprimel = -(a * a + cleft * cleft) / (2 * a * a * a);
        else vprimel = -(1.0 / cleft) * pow(pstar / pleft, -gf);
        if (pstar >= pright) vprimer = (b * b + cright * cright) / (2 * b * b * b);
        else vprimer = (1.0 / cright) * pow(pstar / pright, -gf);
        double vstarl = vxleft - (pstar - pleft) / a;
        double vstarr = vxright + (pstar - pright) / b;
        pstar = pstar - (vstarl - vstarr) / (vprimel - vprimer);
        if (pstar <= 0) pstar = pstarold / 1000.0;
        vstar = (pleft - pright + a * vxleft + b * vxright) / (a + b);
        err = fabs(pstarold - pstar) / pstarold;
    }
    return nnewt;
}

struct Flux { double p, r, u, v, w; };   // pm, rhom, vxm, vym, vzm at an interface

// riemannsolver.F90:32-285 for interface (L | R).  The high-density-peak
// pressure floor (:279-285) is applied by the caller, which knows the
// stencil.  nit receives the Newton iteration count (0 if none ran).
template <class R>
__device__ __forceinline__ Flux riemann_solve(const RScal& Rs, const Face<R>& L, const Face<R>& Rt, int& nit)
{
    const double gamma = Rs.gamma, m = Rs.m, rm = Rs.rm, rgamma = Rs.rgamma, rgamma1 = Rs.rgamma1;
    const double pleft = L.p, pright = Rt.p, rholeft = L.r, rhoright = Rt.r;
    const double vxleft = L.u, vxright = Rt.u;
    double pm = 0.0, rhom = 0.0, vxm = 0.0;
    nit = 0;
    const double cleft = sqrt(gamma * pleft / rholeft);
    const double cright = sqrt(gamma * pright / rhoright);
    double p1, p2, rho1, c1, c2;
    if (pleft <= pright) { p1 = pleft; p2 = pright; rho1 = rholeft; c1 = cleft; c2 = cright; }
    else { p1 = pright; p2 = pleft; rho1 = rhoright; c1 = cright; c2 = cleft; }
    const double dvv = vxleft - vxright;
    const double dv2s = (p2 - p1) / sqrt(0.50 * rho1 * ((gamma - 1) * p1 + (gamma + 1) * p2));
    const double dv1s = -(2.0 * c2 / (gamma - 1.0)) * (1.0 - pow(p1 / p2, m));
    const double dv2r = -2.0 * (c1 + c2) * rgamma1;

    if (dvv >= dv2s) {                                   // two shocks
        pm = tmax(pleft, pright);
        nit = newton(Rs, pm, vxm, pleft, pright, rholeft, rhoright, vxleft, vxright);
        double rs1 = rhoshock(Rs, pm, pleft, rholeft);
        double rs2 = rhoshock(Rs, pm, pright, rhoright);
        double vs1 = (rs1 == rholeft) ? 0.0 : (rs1 * vxm - rholeft * vxleft) / (rs1 - rholeft);
        double vs2 = (rs2 == rhoright) ? 0.0 : (rs2 * vxm - rhoright * vxright) / (rs2 - rhoright);
        if (vs2 < 0.0) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vxm <= 0.0 && vs2 >= 0.0) rhom = rs2;
        else if (vs1 <= 0.0 && vxm > 0.0) rhom = rs1;
        else { pm = pleft; vxm = vxleft; rhom = rholeft; }
    } else if (dvv < dv2s && dvv >= dv1s) {              // shock + rarefaction
        pm = pow((cleft + cright + (vxleft - vxright) * (gamma - 1) * 0.5) /
                     (cleft / pow(pleft, m) + cright / pow(pright, m)), rm);
        nit = newton(Rs, pm, vxm, pleft, pright, rholeft, rhoright, vxleft, vxright);
        if (pleft >= pright) {
            double rs2 = rhoshock(Rs, pm, pright, rhoright);
            double rc1 = rholeft * pow(pm / pleft, rgamma);
            double vs2 = (rs2 == rhoright) ? 0.0 : (rs2 * vxm - rhoright * vxright) / (rs2 - rhoright);
            double vtail1 = vxm - sqrt(gamma * pm / rc1);
            double vhead1 = vxleft - cleft;
            if (vs2 < 0) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vhead1 >= 0.0) { pm = pleft; vxm = vxleft; rhom = rholeft; }
            else if (vxm <= 0.0 && vs2 >= 0) rhom = rs2;
            else if (vxm > 0.0 && vtail1 <= 0.0) rhom = rc1;
            else if (vtail1 > 0.0 && vhead1 <= 0.0) {
                pm = rarwave(pleft, pm, vtail1, vhead1);
                rhom = rarwave(rholeft, rc1, vtail1, vhead1);
                vxm = rarwave(vxleft, vxm, vtail1, vhead1);
            }
        } else {
            double rs1 = rhoshock(Rs, pm, pleft, rholeft);
            double rc2 = rhoright * pow(pm / pright, rgamma);
            double vs1 = (rs1 == rholeft) ? 0.0 : (rs1 * vxm - rholeft * vxleft) / (rs1 - rholeft);
            double vtail2 = vxm + sqrt(gamma * pm / rc2);
            double vhead2 = vxright + cright;
            if (vs1 > 0) { pm = pleft; vxm = vxleft; rhom = rholeft; }
            else if (vhead2 <= 0.0) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vxm >= 0.0 && vs1 <= 0.0) rhom = rs1;
            else if (vxm < 0.0 && vtail2 >= 0.0) rhom = rc2;
            else if (vtail2 < 0.0 && vhead2 >= 0.0) {
                pm = rarwave(pright, pm, vtail2, vhead2);
                rhom = rarwave(rhoright, rc2, vtail2, vhead2);
                vxm = rarwave(vxright, vxm, vtail2, vhead2);
            }
        }
    } else if (dvv < dv1s && dvv > dv2r) {               // two rarefactions
        double zeta = (cright / cleft) * pow(pleft / pright, m);
        double vtleft = vxleft + 2.0 * cleft * rgamma1;
        double vtright = vxright - 2.0 * cright * rgamma1;
        vxm = (vtleft * zeta + vtright) / (1.0 + zeta);
        pm = pleft * pow((gamma - 1) * (vtleft - vtright) / (2.0 * cleft * (1.0 + zeta)), rm);
        double rc1 = rholeft * pow(pm / pleft, rgamma);
        double rc2 = rhoright * pow(pm / pright, rgamma);
        double vtail2 = vxm + sqrt(gamma * pm / rc2);
        double vhead2 = vxright + cright;
        double vtail1 = vxm - sqrt(gamma * pm / rc1);
        double vhead1 = vxleft - cleft;
        if (vhead2 <= 0.0) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vhead1 >= 0.0) { pm = pleft; vxm = vxleft; rhom = rholeft; }
        else if (vhead2 > 0.0 && vtail2 <= 0.0) {
            pm = rarwave(pright, pm, vtail2, vhead2);
            rhom = rarwave(rhoright, rc2, vtail2, vhead2);
            vxm = rarwave(vxright, vxm, vtail2, vhead2);
        } else if (vxm <= 0.0 && vtail2 > 0.0) rhom = rc2;
        else if (vxm > 0.0 && vtail1 <= 0.0) rhom = rc1;
        else if (vtail1 > 0.0 && vhead1 <= 0.0) {
            pm = rarwave(pleft, pm, vtail1, vhead1);
            rhom = rarwave(rholeft, rc1, vtail1, vhead1);
            vxm = rarwave(vxleft, vxm, vtail1, vhead1);
        }
    } else if (dvv <= dv2r) {                            // vacuum
        pm = 0.0;
        vxm = 0.0;
        double vtail2 = vxright - 2.0 * cright * rgamma1;
        double vhead2 = vxright + cright;
        double vtail1 = vxleft - 2.0 * cleft * rgamma1;
        double vhead1 = vxleft - cleft;
        if (vhead2 <= 0.0) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vhead1 >= 0.0) { pm = pleft; vxm = vxleft; rhom = rholeft; }
        else if (vhead2 > 0.0 && vtail2 <= 0.0) {
            pm = rarwave(pright, pm, vtail2, vhead2);
            rhom = rarwave(rhoright, 0.0, vtail2, vhead2);
            vxm = rarwave(vxright, vxm, vtail2, vhead2);
        } else if (vtail1 <= 0.0 && vtail2 > 0.0) rhom = 0.0;
        else if (vtail1 > 0.0 && vhead1 <= 0.0) {
            pm = rarwave(pleft, pm, vtail1, vhead1);
            rhom = rarwave(rholeft, 0.0, vtail1, vhead1);
            vxm = rarwave(vxleft, vxm, vtail1, vhead1);
        }
    }
    Flux F;
    F.p = pm;
    F.r = rhom;
    F.u = vxm;
    if (vxm >= 0.0) { F.v = Rs.dadt * (double)L.v; F.w = Rs.dadt * (double)L.w; }
    else { F.v = Rs.dadt * (double)Rt.v; F.w = Rs.dadt * (double)Rt.w; }
    return F;
}

// High-density-peak pressure floor, riemannsolver.F90:279-285.  rb1..rb4 =
// rho of cells i-3..i of interface i-1/2, u2/u3 the normal velocity of cells
// i-2, i-1.
__device__ __forceinline__ void peak_floor(const RScal& R, Flux& F, double rb1, double rb2, double rb3,
                                           double rb4, double u2, double u3)
{
    if (rb2 > R.dmax * rb1 && rb3 > R.dmax * rb4) {
        double maxp = (u2 - u3) * tmax(rb2, rb3) * tmax(0.0, u2 - u3);
        F.p = tmax(F.p, maxp);
    }
}

// ---------------------------------------------------------------------------
// Conservative update + dual energy, hydro/integrate.F90:1-113 (FP64).
// F1/F2 are the fluxes at the cell's left/right interfaces; the cell's state
// is (pres, rho, u, v, w) in sweep order; etot is the total energy built from
// the physical components in x,y,z order (hydro/total.F90:15).  pmean is
// (pm1+pm2)*0.5 as in hydro/ppm.F90:583 (undefined in integrate.F90).
struct IScal {
    double rdtath, rdx, gamma, rgamma1, eta1, eta2;
};

// Dual-energy selection, integrate.F90:196-216: eintnew enters as the
// internal-energy-equation value and leaves as the selected one; pnew is the
// new pressure.  eint/pres are the cell's old internal energy and pressure.
__host__ __device__ __forceinline__ void dual_energy(const IScal& S, double etotnew, double rhonew, double vxnew,
                                                     double vynew, double vznew, double eint, double pres,
                                                     double& eintnew, double& pnew)
{
    const double gamma = S.gamma;
    double vvv = vxnew * vxnew + vynew * vynew + vznew * vznew;
    double eeaux = etotnew - 0.50 * rhonew * vvv;
    double ee1 = eeaux / etotnew;
    if (ee1 >= S.eta2) eintnew = eeaux;
    if (ee1 >= S.eta1) pnew = (gamma - 1) * eeaux;
    else pnew = (gamma - 1) * eintnew;
    if (eintnew <= 0.0) { eintnew = eint; pnew = pres; }
}

// integrate() with the caller's total and internal energy (integrate.F90:1-2:
// ppmsolver passes etot = total(...) and eint = internal(p), :149-150).
__device__ __forceinline__ void integrate_cell_e(const IScal& S, const Flux& F1, const Flux& F2, double& pres,
                                                 double& rho, double& vx, double& vy, double& vz,
                                                 double etot, double eint, double ghalf, bool grav)
{
    const double rdtath = S.rdtath, rdx = S.rdx, gamma = S.gamma, rgamma1 = S.rgamma1;
    double flux1, flux2;
    flux1 = F1.r * F1.u;
    flux2 = F2.r * F2.u;
    double rhonew = rho + rdtath * (flux1 - flux2) * rdx;
    flux1 = F1.r * F1.u * F1.u + F1.p;
    flux2 = F2.r * F2.u * F2.u + F2.p;
    double vxnew;
    if (grav) {
        double g1 = (rho + rhonew) * ghalf * 0.5;
        vxnew = (vx * rho + rdtath * (flux1 - flux2) * rdx + rdtath * g1) / rhonew;
    } else {
        vxnew = (vx * rho + rdtath * (flux1 - flux2) * rdx) / rhonew;
    }
    flux1 = F1.v * F1.u * F1.r;
    flux2 = F2.v * F2.u * F2.r;
    double vynew = (rho * vy + rdtath * (flux1 - flux2) * rdx) / rhonew;
    flux1 = F1.w * F1.u * F1.r;
    flux2 = F2.w * F2.u * F2.r;
    double vznew = (rho * vz + rdtath * (flux1 - flux2) * rdx) / rhonew;
    flux1 = (gamma * F1.p * rgamma1 + 0.50 * F1.r * (F1.u * F1.u + F1.v * F1.v + F1.w * F1.w)) * F1.u;
    flux2 = (gamma * F2.p * rgamma1 + 0.50 * F2.r * (F2.u * F2.u + F2.v * F2.v + F2.w * F2.w)) * F2.u;
    double etotnew;
    if (grav) {
        double g2 = (rho * vx + rhonew * vxnew) * ghalf * 0.5;
        etotnew = etot + rdtath * (flux1 - flux2) * rdx + rdtath * g2;
    } else {
        etotnew = etot + rdtath * (flux1 - flux2) * rdx;
    }
    double pmean = (F1.p + F2.p) * 0.5;
    double eintnew = eint + rdtath * pmean * (F1.u - F2.u) * rdx;
    flux1 = (F1.p * F1.u) * rgamma1;
    flux2 = (F2.p * F2.u) * rgamma1;
    eintnew = eintnew + rdtath * (flux1 - flux2) * rdx;
    if (eintnew < 0) eintnew = eint;
    double pnew;
    dual_energy(S, etotnew, rhonew, vxnew, vynew, vznew, eint, pres, eintnew, pnew);
    if (rhonew <= 0.) rhonew = rho;
    pres = pnew;
    rho = rhonew;
    vx = vxnew;
    vy = vynew;
    vz = vznew;
}

// The sweep's form: eint = internal(p) = p/(gamma-1) (internal.F90:14).
__device__ __forceinline__ void integrate_cell(const IScal& S, const Flux& F1, const Flux& F2, double& pres,
                                               double& rho, double& vx, double& vy, double& vz,
                                               double etot, double ghalf, bool grav)
{
    integrate_cell_e(S, F1, F2, pres, rho, vx, vy, vz, etot, pres * S.rgamma1, ghalf, grav);
}

}  // namespace hp
