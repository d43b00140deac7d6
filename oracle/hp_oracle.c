/*
 * hp_oracle.c -- CPU FP64 restatement of Hydropad's STENCIL PPM hydro path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels: it
 * may be loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, never by the product path.
 *
 * It restates the reference routines one by one, in the reference's own
 * structure (per-interface gather, two-pass order2interp, three separate
 * loops A/B/C per sweep with full-size temporaries) and in the reference's
 * own order of floating-point operations, so that, built with
 * -O2 -ffp-contract=off on x86-64 without FMA, it reproduces the
 * reference's FP64 results bit for bit (pow/sqrt/div from glibc, as the
 * flang build of the reference uses).  The reference's default-REAL (FP32)
 * literals inside REAL(8) expressions are reproduced as (double)<x>f.
 *
 * Pinning (tests/test_oracle_*.py):
 *   - against vectors of the reference itself, built from its Fortran sources
 *     by oracle/Makefile into oracle/_ref (and frozen under tests/golden/);
 *   - against closed forms: the exact Sod star state, Rankine-Hugoniot,
 *     linear/constant profiles, uniform-state invariance, conservation.
 * Not pinned by the reference: integrate's `pmean`, an uninitialised local in
 * the reference's active build (hydro/integrate.F90:19,60-75; its flang build
 * returns a NaN pressure in that regime).  It is taken as (pm1+pm2)*0.5 from
 * hydro/ppm.F90:583 and pinned by convergence (tests/test_oracle_dual_energy.py):
 * with every cell on the internal-energy equation (e/E < eta1) a smooth wave
 * must approach the total-energy run at second order, which a one-sided,
 * doubled or sign-flipped pmean (planted with or_plant) fails.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/hydropad_gpu.h"

/* FP32 literals of the reference, widened to double (Fortran default REAL). */
#define F_SIXTH   ((double)0.16666667f)   /* hydro/ppm_functs.F90:33-34, hydro/dda.F90:198-199 */
#define F_THIRD   ((double)0.33333333f)   /* hydro/ppm_functs.F90:103,122 */
#define F_0010    ((double)0.010f)        /* hydro/dda.F90:200 */
#define F_0050    ((double)0.050f)        /* hydro/dda.F90:203 */
#define F_1EM15   ((double)1e-15f)        /* hydro/order2interp.F90:176-189 */
#define F_TERR    ((double)1e-3f)         /* hydro/riemann.F90:48 */

#define IDX(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(k)))

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax2(double a, double b) { return a > b ? a : b; }

/* ---------------------------------------------------------------------- */
/* derived constants: hydropad/initialize_parameters.F90:23-24,64-69,183-203 */
int or_params_init(double gamma, int ngrid_max, double dt, hp_params* o)
{
    memset(o, 0, sizeof(*o));
    o->gamma = gamma;
    o->rgamma = 1.0 / gamma;
    o->gamma1 = (gamma + 1.0) / (gamma - 1.0);
    o->rgamma1 = 1.0 / (gamma - 1);
    o->m = (gamma - 1.0) / (2.0 * gamma);
    o->rm = 1.0 / o->m;
    o->gf = (gamma + 1.0) / (2.0 * gamma);
    o->dx = (double)(1.0f / (float)ngrid_max);      /* FP32 division, :23 */
    o->rdx = 1.0 / o->dx;
    o->dt = dt;
    o->at = 1.0;
    o->dat = 0.0;
    o->rat = 1.0 / o->at;
    o->rdtath = dt / 1.0;                            /* ath = 1 */
    o->flatvalue = 0.15;                             /* inputfiles/input.nml:22-23 */
    o->mingradflat = 0.25;
    o->eta1 = 1e-3;                                  /* input.nml:30-31 (REAL(8)) */
    o->eta2 = 0.1;
    o->dmax = 2.0e10;
    return 0;
}

/* ---------------------------------------------------------------------- */
/* hydro/ppm_functs.F90:167-181 */
static double ajside(double am1, double a, double ap1)
{
    if ((ap1 - a) * (a - am1) > 0.0) {
        double da = 0.50 * (ap1 - am1);
        double mn = dmin(dmin(fabs(da), 2.0 * fabs(a - am1)), 2.0 * fabs(ap1 - a));
        return mn * da / fabs(da);
    }
    return 0.0;
}

/* hydro/dda.F90:182-216 (contact steepening) */
static void dda(const hp_params* P, const double* a, double* eta, double* ajld, double* ajrd,
                double dmalf, double dmarg)
{
    const double dx = P->dx;
    double deltam2 = F_SIXTH * (a[2] - 2.0 * a[1] + a[0]) / (dx * dx);
    double deltap2 = F_SIXTH * (a[4] - 2.0 * a[3] + a[2]) / (dx * dx);
    double agrad = fabs(a[3] - a[1]) - F_0010 * dmin(fabs(a[3]), fabs(a[1]));
    if (-deltam2 * deltap2 > 0.0 && agrad > 0.0) {
        double e = -(deltap2 - deltam2) * dx * dx / (a[3] - a[1]);
        *eta = dmax2(0.0, dmin(20.0 * (e - F_0050), 1.0));
    } else {
        *eta = 0.0;
    }
    *ajld = a[1] + 0.50 * dmalf;
    *ajrd = a[3] - 0.50 * dmarg;
}

/* hydro/ppm_functs.F90:91-105 */
static double fl(const double* ajr, const double* d1, const double* a6, double y, double dx)
{
    double x = y / dx;
    return ajr[0] - 0.50 * x * (d1[0] - (1.0 - 2.0 * x * F_THIRD) * a6[0]);
}

/* hydro/ppm_functs.F90:110-124 */
static double fr(const double* ajl, const double* d1, const double* a6, double y, double dx)
{
    double x = y / dx;
    return ajl[1] + 0.50 * x * (d1[1] + (1.0 - 2.0 * x * F_THIRD) * a6[1]);
}

/* hydro/ppm_functs.F90:129-144 */
static double beta(const hp_params* P, double ct, double ut, double uc, double pt, double pc,
                   double gc, double sign)
{
    double gvc = (-gc + P->dat * uc) * P->rat;
    return -sign * ((ut - uc) + sign * (pt - pc) / ct + sign * P->dt * (2.0 * P->dat * pc / (ct * P->at))
                    + P->dt * gvc) / (2 * ct);
}

/* hydro/ppm_functs.F90:149-162 */
static double beta0(const hp_params* P, double ct, double rhot, double rhoc, double pt, double pc)
{
    return (pt - pc) / (ct * ct) + 1.0 / rhot - 1.0 / rhoc + (2.0 * P->dat * pc) / (P->at * ct * ct) * P->dt;
}

/* hydro/ppm_functs.F90:1-86.  a[0..4] is the 5-cell window centred on the
 * cell whose parabola goes to slot 2; slot 1 holds the previous cell's. */
static void correction(const hp_params* P, double addp, double addrho, const double* a,
                       double* ajl, double* ajr, double* deltaa1, double* a6, const double* ccc,
                       double vi, double vip1, double* atl, double* atr, double* acl, double* acr,
                       double flat, int ndda)
{
    const double dt = P->dt, rat = P->rat, dx = P->dx;
    double mmmax = 0.0, y, cv, eta, ajld, ajrd;
    double dmaj = ajside(a[0], a[1], a[2]);
    double dmajp1 = ajside(a[1], a[2], a[3]);
    double dmajp2 = ajside(a[2], a[3], a[4]);
    ajl[1] = a[1] + 0.50 * (a[2] - a[1]) + (dmaj - dmajp1) * F_SIXTH;
    ajr[1] = a[2] + 0.50 * (a[3] - a[2]) + (dmajp1 - dmajp2) * F_SIXTH;
    if (ndda == 1 && addrho >= addp) {
        dda(P, a, &eta, &ajld, &ajrd, dmaj, dmajp2);
        ajl[1] = (1.0 - eta) * ajl[1] + eta * ajld;
        ajr[1] = (1.0 - eta) * ajr[1] + eta * ajrd;
    }
    deltaa1[1] = ajr[1] - ajl[1];
    a6[1] = 6 * (a[2] - 0.50 * (ajl[1] + ajr[1]));
    if ((ajr[1] - a[2]) * (a[2] - ajl[1]) <= 0) {
        ajl[1] = a[2];
        ajr[1] = a[2];
    } else if (deltaa1[1] * a6[1] > deltaa1[1] * deltaa1[1]) {
        ajl[1] = 3.0 * a[2] - 2.0 * ajr[1];
    } else if (deltaa1[1] * a6[1] < (-(deltaa1[1] * deltaa1[1]))) {
        ajr[1] = 3.0 * a[2] - 2.0 * ajl[1];
    }
    ajl[1] = flat * a[2] + (1.0 - flat) * ajl[1];
    ajr[1] = flat * a[2] + (1.0 - flat) * ajr[1];
    deltaa1[1] = ajr[1] - ajl[1];
    a6[1] = 6 * (a[2] - 0.50 * (ajl[1] + ajr[1]));
    if (!atl) return;   /* pass i=nmin: only the slot-2 parabola is used later */

    cv = dt * (vi + ccc[0]) * rat;
    y = dmax2(cv, mmmax);
    *atl = fl(ajr, deltaa1, a6, y, dx);
    cv = -dt * (vip1 - ccc[1]) * rat;
    y = dmax2(cv, mmmax);
    *atr = fr(ajl, deltaa1, a6, y, dx);
    y = dt * (vi - ccc[0]) * rat;
    acl[0] = fl(ajr, deltaa1, a6, y, dx);
    y = dt * vi * rat;
    acl[1] = fl(ajr, deltaa1, a6, y, dx);
    y = dt * (vi + ccc[0]) * rat;
    acl[2] = fl(ajr, deltaa1, a6, y, dx);
    y = -dt * (vip1 - ccc[1]) * rat;
    acr[0] = fr(ajl, deltaa1, a6, y, dx);
    y = -dt * vip1 * rat;
    acr[1] = fr(ajl, deltaa1, a6, y, dx);
    y = -dt * (vip1 + ccc[1]) * rat;
    acr[2] = fr(ajl, deltaa1, a6, y, dx);
}

/* flattening test over 4 consecutive pairs, hydro/order2interp.F90:65-113.
 * A pair whose product is 0 is skipped: the reference max-es in a stale
 * ratio there that was already tested against mingradflat. */
static int flat_var(const double* a, int start, double thr)
{
    double gradmax = 0.0;
    for (int ii = 1; ii <= 4; ++ii) {
        int ic = start + ii;
        if (a[ic] * a[ic - 1] != 0.0) {
            double gl = fabs((a[ic] - a[ic - 1]) / dmin(a[ic], a[ic - 1]));
            gradmax = dmax2(gradmax, gl);
        }
    }
    return gradmax >= thr;
}

/* hydro/order2interp.F90:1-265 with ngrid=6, nmin=3, nmax=4 (STENCIL).
 * Arrays are the 6-cell stencils (cells i-3..i+2 of interface i-1/2),
 * 0-based here.  out[10] = pleft,pright,rholeft,rhoright,vxleft,vxright,
 * vyleft,vyright,vzleft,vzright.  *nclamp counts the negative-state clamps
 * of the pass that produces the outputs. */
void or_order2interp(const hp_params* P, const double* pres, const double* rho, const double* vx,
                     const double* vy, const double* vz, const double* g, const double* c,
                     double* out, int64_t* nclamp)
{
    double pjl[2], pjr[2], dp[2], p6[2], rjl[2], rjr[2], dr[2], r6[2];
    double xjl[2], xjr[2], dvx[2], x6[2], yjl[2], yjr[2], dvy[2], y6[2];
    double zjl[2], zjr[2], dvz[2], z6[2], gjl[2], gjr[2], dg[2], g6[2];
    double pcl[3], pcr[3], rcl[3], rcr[3], xcl[3], xcr[3], ycl[3], ycr[3], zcl[3], zcr[3], gcl[3], gcr[3];
    double ptl = 0, ptr = 0, rtl = 0, rtr = 0, xtl = 0, xtr = 0, ytl = 0, ytr = 0, ztl = 0, ztr = 0, gtl, gtr;
    const double gamma = P->gamma, rat = P->rat;

    for (int i = 2; i <= 3; ++i) {        /* Fortran i = nmin..nmax = 3..4 */
        int mi = i - 2;                   /* mainindex - 1 */
        double flat = 0.0;
        double cbar[2] = {c[i - 1], c[i]};
        if (flat_var(pres, mi, P->mingradflat)) flat = P->flatvalue;
        if (flat_var(rho, mi, P->mingradflat)) flat = P->flatvalue;
        if (flat_var(vx, mi, P->mingradflat)) flat = P->flatvalue;
        if (flat_var(vy, mi, P->mingradflat)) flat = P->flatvalue;
        if (flat_var(vz, mi, P->mingradflat)) flat = P->flatvalue;
        /* slot 2 -> slot 1 (order2interp.F90:120-147) */
        pjl[0] = pjl[1]; pjr[0] = pjr[1]; dp[0] = dp[1]; p6[0] = p6[1];
        rjl[0] = rjl[1]; rjr[0] = rjr[1]; dr[0] = dr[1]; r6[0] = r6[1];
        xjl[0] = xjl[1]; xjr[0] = xjr[1]; dvx[0] = dvx[1]; x6[0] = x6[1];
        yjl[0] = yjl[1]; yjr[0] = yjr[1]; dvy[0] = dvy[1]; y6[0] = y6[1];
        zjl[0] = zjl[1]; zjr[0] = zjr[1]; dvz[0] = dvz[1]; z6[0] = z6[1];
        gjl[0] = gjl[1]; gjr[0] = gjr[1]; dg[0] = dg[1]; g6[0] = g6[1];

        double addp = fabs(pres[i + 1] - pres[i - 1]) / dmin(pres[i - 1], pres[i + 1]);
        double addrho = 0.50 * gamma * fabs(rho[i + 1] - rho[i - 1]) / dmin(rho[i - 1], rho[i + 1]);
        int last = (i == 3);
        correction(P, addp, addrho, pres + mi, pjl, pjr, dp, p6, cbar, vx[i - 1], vx[i],
                   last ? &ptl : 0, &ptr, pcl, pcr, flat, 0);
        correction(P, addp, addrho, rho + mi, rjl, rjr, dr, r6, cbar, vx[i - 1], vx[i],
                   last ? &rtl : 0, &rtr, rcl, rcr, flat, 1);
        correction(P, addp, addrho, vx + mi, xjl, xjr, dvx, x6, cbar, vx[i - 1], vx[i],
                   last ? &xtl : 0, &xtr, xcl, xcr, flat, 0);
        correction(P, addp, addrho, vy + mi, yjl, yjr, dvy, y6, cbar, vx[i - 1], vx[i],
                   last ? &ytl : 0, &ytr, ycl, ycr, flat, 0);
        correction(P, addp, addrho, vz + mi, zjl, zjr, dvz, z6, cbar, vx[i - 1], vx[i],
                   last ? &ztl : 0, &ztr, zcl, zcr, flat, 0);
        correction(P, addp, addrho, g + mi, gjl, gjr, dg, g6, cbar, vx[i - 1], vx[i],
                   last ? &gtl : 0, &gtr, gcl, gcr, flat, 0);
        if (!last) continue;

        int64_t ncl = 0;
        if (ptl < 0.0) { ptl = F_1EM15; ++ncl; }
        if (ptr < 0.0) { ptr = F_1EM15; ++ncl; }
        if (rtl < 0.0) { rtl = F_1EM15; ++ncl; }
        if (rtr < 0.0) { rtr = F_1EM15; ++ncl; }
        if (nclamp) *nclamp += ncl;

        double ctl = sqrt(gamma * ptl * rtl);
        double ctr = sqrt(gamma * ptr * rtr);
        double betaml, betapl, beta0l, betamr, betapr, beta0r, vyleft, vzleft, vyright, vzright;
        double lambdam = (vx[i - 1] - cbar[0]) * rat;
        double lambda0 = vx[i - 1] * rat;
        double lambdap = (vx[i - 1] + cbar[0]) * rat;
        if (lambdam <= 0.0) betaml = 0.0;
        else betaml = beta(P, ctl, xtl, xcl[0], ptl, pcl[0], gcl[0], -1.0);
        if (lambdap <= 0.0) { betapl = 0.0; vyleft = ytl; vzleft = ztl; }
        else { betapl = beta(P, ctl, xtl, xcl[2], ptl, pcl[2], gcl[2], 1.0); vyleft = ytl; vzleft = ztl; }
        if (lambda0 <= 0.0) beta0l = 0.0;
        else { beta0l = beta0(P, ctl, rtl, rcl[1], ptl, pcl[1]); vyleft = ycl[1]; vzleft = zcl[1]; }

        lambdam = (vx[i] - cbar[1]) * rat;
        lambda0 = vx[i] * rat;
        lambdap = (vx[i] + cbar[1]) * rat;
        if (lambdap >= 0.0) betapr = 0.0;
        else betapr = beta(P, ctr, xtr, xcr[2], ptr, pcr[2], gcr[2], 1.0);
        if (lambdam >= 0.0) { betamr = 0.0; vyright = ytr; vzright = ztr; }
        else { betamr = beta(P, ctr, xtr, xcr[0], ptr, pcr[0], gcr[0], -1.0); vyright = ytr; vzright = ztr; }
        if (lambda0 >= 0.0) beta0r = 0.0;
        else { beta0r = beta0(P, ctr, rtr, rcr[1], ptr, pcr[1]); vyright = ycr[1]; vzright = zcr[1]; }

        double pleft = ptl + ctl * ctl * (betapl + betaml);
        if (pleft <= 0.0) pleft = ptl;
        double pright = ptr + ctr * ctr * (betapr + betamr);
        if (pright <= 0.0) pright = ptr;
        double vxleft = xtl + ctl * (betapl - betaml);
        double vxright = xtr + ctr * (betapr - betamr);
        double rholeft = 1.0 / (1.0 / rtl - betaml - beta0l - betapl);
        if (rholeft <= 0.0) rholeft = rtl;
        double rhoright = 1.0 / (1.0 / rtr - betamr - beta0r - betapr);
        if (rhoright <= 0.0) rhoright = rtr;
        out[0] = pleft; out[1] = pright; out[2] = rholeft; out[3] = rhoright;
        out[4] = vxleft; out[5] = vxright; out[6] = vyleft; out[7] = vyright;
        out[8] = vzleft; out[9] = vzright;
    }
    (void)gtl; (void)gtr;
}

/* ---------------------------------------------------------------------- */
/* hydro/riemann.F90:1-9 */
static double rarwave(double f1, double f2, double vt, double vh)
{
    return f2 + vt * (f1 - f2) / (vt - vh);
}

/* hydro/riemann.F90:14-24 */
static double rhoshock(const hp_params* P, double p2, double p1, double rho1)
{
    return rho1 * (P->gamma1 * p2 / p1 + 1.0) / (P->gamma1 + p2 / p1);
}

/* The reference's nnewtmax = 100 (hydro/riemann.F90:51); settable so that
 * tests can reach the non-convergence path with ordinary states. */
static int newton_max = 100;
void or_set_newton_max(int n) { newton_max = n > 0 ? n : 100; }

/* hydro/riemann.F90:29-111 (van Leer Newton iteration).  Returns the number
 * of iterations, or -1 where the reference would `stop` (>nnewtmax). */
int or_riemann(const hp_params* P, double* pstar, double* vstar, double pleft, double pright,
               double rholeft, double rhoright, double vxleft, double vxright)
{
    const double gamma = P->gamma, m = P->m, gf = P->gf;
    double err = 1.e10, terr = F_TERR, a = 0, b = 0;
    int nnewt = 0, nnewtmax = newton_max;
    double kappa = sqrt(gamma * (pleft + pright) * (rholeft + rhoright) * 0.25);
    double pcheck = 0.50 * (pleft + pright) + 0.50 * kappa * (vxleft + vxright);
    if (pcheck > 0) *pstar = pcheck;
    while (err > terr) {
        nnewt = nnewt + 1;
        if (nnewt > nnewtmax) return -1;
        double pstarold = *pstar;
        double cleft = sqrt(gamma * pleft * rholeft);
        double cright = sqrt(gamma * pright * rhoright);
        double pstup = (1. - pow(*pstar / pleft, m));
        if (*pstar >= pleft) a = cleft * sqrt(gf * *pstar / pleft + m);
        else if (pstup == 0.) a = cleft * sqrt(gf * *pstar / pleft + m);
        else a = m * cleft * (1.0 - *pstar / pleft) / (1 - pow(*pstar / pleft, m));
        pstup = (1. - pow(*pstar / pright, m));
        if (*pstar >= pright) b = cright * sqrt(gf * *pstar / pright + m);
        else if (pstup == 0.) b = cright * sqrt(gf * *pstar / pright + m);
        else b = m * cright * (1.0 - *pstar / pright) / (1 - pow(*pstar / pright, m));
        double vprimel, vprimer;
        if (*pstar >= pleft) vprimel = -(a * a + cleft * cleft) / (2 * a * a * a);
        else vprimel = -(1.0 / cleft) * pow(*pstar / pleft, -gf);
        if (*pstar >= pright) vprimer = (b * b + cright * cright) / (2 * b * b * b);
        else vprimer = (1.0 / cright) * pow(*pstar / pright, -gf);
        double vstarl = vxleft - (*pstar - pleft) / a;
        double vstarr = vxright + (*pstar - pright) / b;
        *pstar = *pstar - (vstarl - vstarr) / (vprimel - vprimer);
        if (*pstar <= 0) *pstar = pstarold / 1000.0;
        *vstar = (pleft - pright + a * vxleft + b * vxright) / (a + b);
        err = fabs(pstarold - *pstar) / pstarold;
    }
    return nnewt;
}

/* hydro/riemannsolver.F90:1-289.  in[10] as or_order2interp's out; rhobar,
 * vxbar are the 6-cell stencils of the interface (dmax check, :279-285).
 * out[5] = pm, rhom, vxm, vym, vzm.  Returns Newton iterations (0 if the
 * branch needs none), -1 on non-convergence. */
int or_riemannsolver(const hp_params* P, const double* rhobar, const double* vxbar, const double* in,
                     double* out)
{
    const double gamma = P->gamma, m = P->m, rm = P->rm, rgamma = P->rgamma, rgamma1 = P->rgamma1;
    double pleft = in[0], pright = in[1], rholeft = in[2], rhoright = in[3];
    double vxleft = in[4], vxright = in[5], vyleft = in[6], vyright = in[7], vzleft = in[8], vzright = in[9];
    double pm = 0, rhom = 0, vxm = 0, vym, vzm;
    double p1, p2, rho1, rho2, c1, c2;
    int nit = 0;
    double cleft = sqrt(gamma * pleft / rholeft);
    double cright = sqrt(gamma * pright / rhoright);
    double dadt = 1.0 - 0.5 * P->dat * P->dt * P->rat;
    double mmmax = 0.0;
    if (pleft <= pright) { p1 = pleft; p2 = pright; rho1 = rholeft; rho2 = rhoright; c1 = cleft; c2 = cright; }
    else { p1 = pright; p2 = pleft; rho1 = rhoright; rho2 = rholeft; c1 = cright; c2 = cleft; }
    (void)rho2;
    double dvv = vxleft - vxright;
    double dv2s = (p2 - p1) / sqrt(0.50 * rho1 * ((gamma - 1) * p1 + (gamma + 1) * p2));
    double dv1s = -(2.0 * c2 / (gamma - 1.0)) * (1.0 - pow(p1 / p2, m));
    double dv2r = -2.0 * (c1 + c2) * rgamma1;

    if (dvv >= dv2s) {
        pm = dmax2(pleft, pright);
        nit = or_riemann(P, &pm, &vxm, pleft, pright, rholeft, rhoright, vxleft, vxright);
        double rs1 = rhoshock(P, pm, pleft, rholeft);
        double rs2 = rhoshock(P, pm, pright, rhoright);
        double vs1 = (rs1 == rholeft) ? 0.0 : (rs1 * vxm - rholeft * vxleft) / (rs1 - rholeft);
        double vs2 = (rs2 == rhoright) ? 0.0 : (rs2 * vxm - rhoright * vxright) / (rs2 - rhoright);
        if (vs2 < 0.0) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vxm <= 0.0 && vs2 >= 0.0) rhom = rs2;
        else if (vs1 <= 0.0 && vxm > 0.0) rhom = rs1;
        else { pm = pleft; vxm = vxleft; rhom = rholeft; }
    } else if (dvv < dv2s && dvv >= dv1s) {
        pm = pow((cleft + cright + (vxleft - vxright) * (gamma - 1) * 0.5) /
                 (cleft / pow(pleft, m) + cright / pow(pright, m)), rm);
        nit = or_riemann(P, &pm, &vxm, pleft, pright, rholeft, rhoright, vxleft, vxright);
        if (pleft >= pright) {
            double rs2 = rhoshock(P, pm, pright, rhoright);
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
            double rs1 = rhoshock(P, pm, pleft, rholeft);
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
    } else if (dvv < dv1s && dvv > dv2r) {
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
    } else if (dvv <= dv2r) {
        pm = 0.0; vxm = 0.0;
        double rc1 = 0.0, rc2 = 0.0;
        double vtail2 = vxright - 2.0 * cright * rgamma1;
        double vhead2 = vxright + cright;
        double vtail1 = vxleft - 2.0 * cleft * rgamma1;
        double vhead1 = vxleft - cleft;
        if (vhead2 <= 0.0) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vhead1 >= 0.0) { pm = pleft; vxm = vxleft; rhom = rholeft; }
        else if (vhead2 > 0.0 && vtail2 <= 0.0) {
            pm = rarwave(pright, pm, vtail2, vhead2);
            rhom = rarwave(rhoright, rc2, vtail2, vhead2);
            vxm = rarwave(vxright, vxm, vtail2, vhead2);
        } else if (vtail1 <= 0.0 && vtail2 > 0.0) rhom = 0.0;
        else if (vtail1 > 0.0 && vhead1 <= 0.0) {
            pm = rarwave(pleft, pm, vtail1, vhead1);
            rhom = rarwave(rholeft, rc1, vtail1, vhead1);
            vxm = rarwave(vxleft, vxm, vtail1, vhead1);
        }
    }
    if (vxm >= 0.0) { vym = dadt * vyleft; vzm = dadt * vzleft; }
    else { vym = dadt * vyright; vzm = dadt * vzright; }
    if (rhobar[1] > P->dmax * rhobar[0] && rhobar[2] > P->dmax * rhobar[3]) {
        double maxp = (vxbar[1] - vxbar[2]) * dmax2(rhobar[1], rhobar[2]) * dmax2(mmmax, vxbar[1] - vxbar[2]);
        pm = dmax2(pm, maxp);
    }
    out[0] = pm; out[1] = rhom; out[2] = vxm; out[3] = vym; out[4] = vzm;
    return nit;
}

/* ---------------------------------------------------------------------- */
/* hydro/internal.F90:14, hydro/total.F90:15 */
double or_internal(const hp_params* P, double pres) { return pres * P->rgamma1; }
double or_total(const hp_params* P, double pres, double rho, double vx, double vy, double vz)
{
    return pres * P->rgamma1 + 0.5 * rho * (vx * vx + vy * vy + vz * vz);
}

/* Planted mistakes for the tests that show a check can fail (0 = none):
 * 1 pmean = pm1, 2 pmean = pm1 + pm2, 3 pmean = -(pm1 + pm2)/2. */
static int planted = 0;
void or_plant(int what) { planted = what; }

/* hydro/integrate.F90:1-113; pmean from hydro/ppm.F90:583. */
void or_integrate(const hp_params* P, double pm1, double pm2, double rhom1, double rhom2, double vxm1,
                  double vxm2, double vym1, double vym2, double vzm1, double vzm2, double* pres,
                  double* rho, double* vx, double* vy, double* vz, double etot, double eint, double ghalf)
{
    const double rdtath = P->rdtath, rdx = P->rdx, gamma = P->gamma, rgamma1 = P->rgamma1;
    double flux1, flux2, rhonew, vxnew, vynew, vznew, etotnew, eintnew, grav, pnew;
    flux1 = rhom1 * vxm1;
    flux2 = rhom2 * vxm2;
    rhonew = *rho + rdtath * (flux1 - flux2) * rdx;
    flux1 = rhom1 * vxm1 * vxm1 + pm1;
    flux2 = rhom2 * vxm2 * vxm2 + pm2;
    grav = (*rho + rhonew) * ghalf * 0.5;
    vxnew = (*vx * *rho + rdtath * (flux1 - flux2) * rdx + rdtath * grav) / rhonew;
    flux1 = vym1 * vxm1 * rhom1;
    flux2 = vym2 * vxm2 * rhom2;
    vynew = (*rho * *vy + rdtath * (flux1 - flux2) * rdx) / rhonew;
    flux1 = vzm1 * vxm1 * rhom1;
    flux2 = vzm2 * vxm2 * rhom2;
    vznew = (*rho * *vz + rdtath * (flux1 - flux2) * rdx) / rhonew;
    flux1 = (gamma * pm1 * rgamma1 + 0.50 * rhom1 * (vxm1 * vxm1 + vym1 * vym1 + vzm1 * vzm1)) * vxm1;
    flux2 = (gamma * pm2 * rgamma1 + 0.50 * rhom2 * (vxm2 * vxm2 + vym2 * vym2 + vzm2 * vzm2)) * vxm2;
    grav = (*rho * *vx + rhonew * vxnew) * ghalf * 0.5;
    etotnew = etot + rdtath * (flux1 - flux2) * rdx + rdtath * grav;
    double pmean = (pm1 + pm2) * 0.5;
    if (planted == 1) pmean = pm1;
    else if (planted == 2) pmean = pm1 + pm2;
    else if (planted == 3) pmean = -(pm1 + pm2) * 0.5;
    eintnew = eint + rdtath * pmean * (vxm1 - vxm2) * rdx;
    flux1 = (pm1 * vxm1) * rgamma1;
    flux2 = (pm2 * vxm2) * rgamma1;
    eintnew = eintnew + rdtath * (flux1 - flux2) * rdx;
    if (eintnew < 0) eintnew = eint;
    double ee2 = etotnew;
    double vvv = vxnew * vxnew + vynew * vynew + vznew * vznew;
    double eeaux = etotnew - 0.50 * rhonew * vvv;
    double ee1 = eeaux / ee2;
    if (ee1 >= P->eta2) eintnew = eeaux;
    double ee0 = eeaux / etotnew;
    if (ee0 >= P->eta1) pnew = (gamma - 1) * eeaux;
    else pnew = (gamma - 1) * eintnew;
    if (eintnew <= 0.0) { eintnew = eint; pnew = *pres; }
    if (rhonew <= 0.) rhonew = *rho;
    *pres = pnew; *rho = rhonew; *vx = vxnew; *vy = vynew; *vz = vznew;
}

/* ---------------------------------------------------------------------- */
/* Per-direction index maps of hydro/ppmsolver.F90 STENCIL kernels.  For a
 * sweep along axis d, the normal velocity and the transverse pair are
 * (vx;vy,vz) for x (:105-107), (vy;vx,vz) for y (:227-229), (vz;vy,vx) for z
 * (:355-357). */
typedef struct { size_t s_axis, s_a, s_b; int n_axis, n_a, n_b; } sweep_geom;

static int nthreads_req = 1;
void or_set_threads(int n) { nthreads_req = n > 0 ? n : 1; }

/* SUBROUTINE ppmsolver, STENCIL branch (hydro/ppmsolver.F90:54-160 x,
 * :213-288 y, :341-416 z).  Updates the interior of the 5 fields in place.
 * gforce may be NULL (treated as 0).  Returns the largest Newton iteration
 * count, or -1 if any interface did not converge. */
int or_ppmsolver(const hp_params* P, double* p3d, double* rho3d, double* vx3d, double* vy3d, double* vz3d,
                 const double* gforce, int nx, int ny, int nz, int nbound, int ndirection,
                 hp_counters* cnt)
{
    double *vn3d, *vt13d, *vt23d;
    size_t s_axis, s_a, s_b;
    int n_axis, n_a, n_b;
    const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
    switch (ndirection) {
    case 1: vn3d = vx3d; vt13d = vy3d; vt23d = vz3d; s_axis = sx; s_a = sy; s_b = sz; n_axis = nx; n_a = ny; n_b = nz; break;
    case 2: vn3d = vy3d; vt13d = vx3d; vt23d = vz3d; s_axis = sy; s_a = sx; s_b = sz; n_axis = ny; n_a = nx; n_b = nz; break;
    case 3: vn3d = vz3d; vt13d = vy3d; vt23d = vx3d; s_axis = sz; s_a = sx; s_b = sy; n_axis = nz; n_a = nx; n_b = ny; break;
    default: return -2;
    }
    const size_t ntot = (size_t)nx * ny * nz;
    /* the 10 temporaries of hydro/ppmsolver.F90:55-64; B overwrites the *left ones */
    double* T = (double*)malloc(sizeof(double) * ntot * 10);
    if (!T) return -3;
    double *pl = T, *pr = T + ntot, *rl = T + 2 * ntot, *rr = T + 3 * ntot, *xl = T + 4 * ntot,
           *xr = T + 5 * ntot, *yl = T + 6 * ntot, *yr = T + 7 * ntot, *zl = T + 8 * ntot, *zr = T + 9 * ntot;
    const int d = ndirection - 1;
    int64_t nitmax = 0, nonconv = 0, nclamp = 0;
    const int nb = nbound;

    /* kernel A: interfaces i = nb .. n_axis-nb (0-based), interior transverse */
#pragma omp parallel for collapse(2) num_threads(nthreads_req) reduction(+ : nclamp) schedule(static)
    for (int b = nb; b < n_b - nb; ++b)
        for (int a = nb; a < n_a - nb; ++a)
            for (int i = nb; i <= n_axis - nb; ++i) {
                double pres[6], rho[6], vx[6], vy[6], vz[6], g[6], c[6], o[10];
                for (int s = 0; s < 6; ++s) {
                    size_t q = (size_t)(i + s - 3) * s_axis + (size_t)a * s_a + (size_t)b * s_b;
                    pres[s] = p3d[q]; rho[s] = rho3d[q]; vx[s] = vn3d[q]; vy[s] = vt13d[q]; vz[s] = vt23d[q];
                    g[s] = gforce ? -gforce[3 * q + d] : 0.0;
                    c[s] = sqrt(P->gamma * pres[s] / rho[s]);
                }
                or_order2interp(P, pres, rho, vx, vy, vz, g, c, o, &nclamp);
                size_t q = (size_t)i * s_axis + (size_t)a * s_a + (size_t)b * s_b;
                pl[q] = o[0]; pr[q] = o[1]; rl[q] = o[2]; rr[q] = o[3]; xl[q] = o[4];
                xr[q] = o[5]; yl[q] = o[6]; yr[q] = o[7]; zl[q] = o[8]; zr[q] = o[9];
            }
    /* kernel B: Riemann problem per interface */
#pragma omp parallel for collapse(2) num_threads(nthreads_req) reduction(max : nitmax) reduction(+ : nonconv) schedule(static)
    for (int b = nb; b < n_b - nb; ++b)
        for (int a = nb; a < n_a - nb; ++a)
            for (int i = nb; i <= n_axis - nb; ++i) {
                double rhob[6], vxb[6], in[10], o[5];
                for (int s = 0; s < 6; ++s) {
                    /* direction-consistent gather (the y/z reference kernels read
                     * these at transposed indices, ppmsolver.F90:251-252,379-380;
                     * only the dmax branch reads them) */
                    size_t q = (size_t)(i + s - 3) * s_axis + (size_t)a * s_a + (size_t)b * s_b;
                    rhob[s] = rho3d[q]; vxb[s] = vn3d[q];
                }
                size_t q = (size_t)i * s_axis + (size_t)a * s_a + (size_t)b * s_b;
                in[0] = pl[q]; in[1] = pr[q]; in[2] = rl[q]; in[3] = rr[q]; in[4] = xl[q];
                in[5] = xr[q]; in[6] = yl[q]; in[7] = yr[q]; in[8] = zl[q]; in[9] = zr[q];
                int nit = or_riemannsolver(P, rhob, vxb, in, o);
                if (nit < 0) nonconv += 1;
                else if (nit > nitmax) nitmax = nit;
                pl[q] = o[0]; rl[q] = o[1]; xl[q] = o[2]; yl[q] = o[3]; zl[q] = o[4];
            }
    /* kernel C: conservative update + dual energy per interior cell */
#pragma omp parallel for collapse(2) num_threads(nthreads_req) schedule(static)
    for (int b = nb; b < n_b - nb; ++b)
        for (int a = nb; a < n_a - nb; ++a)
            for (int i = nb; i < n_axis - nb; ++i) {
                size_t q = (size_t)i * s_axis + (size_t)a * s_a + (size_t)b * s_b;
                size_t q1 = q + s_axis;
                double etots = or_total(P, p3d[q], rho3d[q], vx3d[q], vy3d[q], vz3d[q]);
                double eints = or_internal(P, p3d[q]);
                double ghalf = gforce ? -gforce[3 * q + d] : 0.0;
                or_integrate(P, pl[q], pl[q1], rl[q], rl[q1], xl[q], xl[q1], yl[q], yl[q1], zl[q], zl[q1],
                             &p3d[q], &rho3d[q], &vn3d[q], &vt13d[q], &vt23d[q], etots, eints, ghalf);
            }
    free(T);
    if (cnt) {
        if (nitmax > cnt->newton_iters_max) cnt->newton_iters_max = nitmax;
        cnt->newton_nonconv += nonconv;
        cnt->neg_state_clamps += nclamp;
    }
    (void)sx;
    return nonconv ? -1 : (int)nitmax;
}

/* Periodic ghost fill of one field, non-MPI branch of
 * hydropad/exchange_mesh.F90:47-123 (z faces, then x, then y).
 * which: 0 = all, 1 = x only, 2 = y only, 3 = z only. */
static void fill_one(double* f, int nx, int ny, int nz, int nb, int which)
{
    if (which == 0 || which == 3)
        for (int k = 0; k < nb; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    f[IDX(i, j, k)] = f[IDX(i, j, nz - 2 * nb + k)];
                    f[IDX(i, j, nz - nb + k)] = f[IDX(i, j, k + nb)];
                }
    if (which == 0 || which == 1)
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    f[IDX(i, j, k)] = f[IDX(nx - 2 * nb + i, j, k)];
                    f[IDX(nx - nb + i, j, k)] = f[IDX(i + nb, j, k)];
                }
    if (which == 0 || which == 2)
        for (int j = 0; j < nb; ++j)
            for (int k = 0; k < nz; ++k)
                for (int i = 0; i < nx; ++i) {
                    f[IDX(i, j, k)] = f[IDX(i, ny - 2 * nb + j, k)];
                    f[IDX(i, ny - nb + j, k)] = f[IDX(i, j + nb, k)];
                }
}

void or_exchange_mesh(double* p, double* rho, double* vx, double* vy, double* vz, int nx, int ny, int nz,
                      int nb, int which)
{
    fill_one(rho, nx, ny, nz, nb, which);
    fill_one(p, nx, ny, nz, nb, which);
    fill_one(vx, nx, ny, nz, nb, which);
    fill_one(vy, nx, ny, nz, nb, which);
    fill_one(vz, nx, ny, nz, nb, which);
}

/* hydropad/speed.F90:17-39 */
double or_speed(const hp_params* P, const double* p3d, const double* rho3d, const double* vx3d,
                const double* vy3d, const double* vz3d, int nx, int ny, int nz, int nb)
{
    double vmax = 0.0;
    for (int k = nb; k < nz - nb; ++k)
        for (int j = nb; j < ny - nb; ++j)
            for (int i = nb; i < nx - nb; ++i) {
                size_t q = IDX(i, j, k);
                double vsound = sqrt(P->gamma * p3d[q] / rho3d[q]);
                vmax = dmax2(vmax, dmax2(dmax2(fabs(vx3d[q]) + vsound, fabs(vy3d[q]) + vsound),
                                         fabs(vz3d[q]) + vsound));
            }
    return vmax;
}

/* hydropad/functions.F90:2-35 (non-cosmological).  Updates P->dt, P->rdtath
 * and *t; returns the new dt. */
double or_timedet(hp_params* P, double velmax, double cour, double tstop, double* t)
{
    double told = *t;
    double rdtc = velmax / (cour * P->dx);
    double dt = (velmax == 0.0) ? 0.1 : 1.0 / rdtc;
    *t = told + dt;
    if (*t > tstop) { *t = tstop; dt = *t - told; }
    P->at = 1.0; P->dat = 0.0; P->rat = 1.0 / P->at;
    P->dt = dt;
    P->rdtath = dt / 1.0;
    return dt;
}

/* Sweep order of hydro/ppm_lxlylz.F90:9-221 */
static const int NROT_ORDER[6][3] = {{1, 2, 3}, {3, 2, 1}, {2, 3, 1}, {1, 3, 2}, {3, 1, 2}, {2, 1, 3}};
int or_nrot_dir(int nrot, int s) { return NROT_ORDER[nrot - 1][s]; }

/* SUBROUTINE ppm_lxlylz: three sweeps, each followed by exchange_mesh. */
int or_ppm_lxlylz(const hp_params* P, double* p, double* rho, double* vx, double* vy, double* vz,
                  const double* gforce, int nx, int ny, int nz, int nb, int nrot, hp_counters* cnt)
{
    int worst = 0;
    for (int s = 0; s < 3; ++s) {
        int r = or_ppmsolver(P, p, rho, vx, vy, vz, gforce, nx, ny, nz, nb, NROT_ORDER[nrot - 1][s], cnt);
        if (r < 0) worst = r;
        or_exchange_mesh(p, rho, vx, vy, vz, nx, ny, nz, nb, 0);
    }
    return worst;
}

/* Time loop of hydropad/main.F90:100-121 and evolve_sys.F90:30-31,54,122-129:
 * nrot = mod(nstep,6)+1 (1 on step 1), ppm_lxlylz, speed, timedet.  P->dt and
 * P->rdtath must hold the dt of step first_step on entry. */
int or_evolve(hp_params* P, double* p, double* rho, double* vx, double* vy, double* vz,
              const double* gforce, int nx, int ny, int nz, int nb, int first_step, int nsteps,
              double cour, double tstop, double* t, hp_counters* cnt)
{
    int worst = 0;
    for (int nstep = first_step; nstep < first_step + nsteps; ++nstep) {
        int nrot = nstep % 6 + 1;
        if (nstep == 1) nrot = 1;
        int r = or_ppm_lxlylz(P, p, rho, vx, vy, vz, gforce, nx, ny, nz, nb, nrot, cnt);
        if (r < 0) worst = r;
        double vmax = or_speed(P, p, rho, vx, vy, vz, nx, ny, nz, nb);
        or_timedet(P, vmax, cour, tstop, t);
    }
    return worst;
}

/* ---- comoving boxes (COSMO, Einstein-de Sitter) ---------------------------
 * FP32 exponents of the reference, widened: 2.0/3.0 and -1.0/3.0 are default
 * REAL quotients (hydropad/functions.F90:67-68,145-148); 0.010 likewise
 * (hydropad/initialize_parameters.F90:141). */
#define F_TWO_THIRDS   ((double)(2.0f / 3.0f))
#define F_MINUS_THIRD  ((double)(-1.0f / 3.0f))

/* hydropad/initialize_parameters.F90:136-178 (COSMO, omega_m = 1, no restart),
 * with at = a_start as set by the initial-conditions reader (initialdm.F90:55). */
void or_cosmo_init(double a_start, double cour, double ca, double tstop, hp_cosmo* s)
{
    double at = a_start;
    double tin = pow(at, 1.5);
    double told = tin;
    double dt = F_0010 * tin;
    double t = tin + dt;
    double thalf = told + dt / 2.0;
    s->cour = cour; s->ca = ca; s->tstop = tstop;
    s->at = pow(told, F_TWO_THIRDS);
    s->atnew = pow(t, F_TWO_THIRDS);
    s->dat = F_TWO_THIRDS * pow(told, F_MINUS_THIRD);
    s->ath = pow(told + dt / 2., F_TWO_THIRDS);
    s->dath = F_TWO_THIRDS * pow(thalf, F_MINUS_THIRD);
    s->datnew = F_TWO_THIRDS * pow(t, F_MINUS_THIRD);
    s->rat = 1.0 / s->at;
    s->rdtath = dt / s->ath;
    s->t = t; s->dt = dt;
    s->nstep = 0;
}

/* SUBROUTINE timedet_cosmo, omega_m = 1 (hydropad/functions.F90:38-151), run
 * at the end of step s->nstep+1.  Not applied: the dt0 = 1e-9 cap (:49,97,
 * which would make any run take ~1e7 steps) and the particle speed (:75-84,
 * uninitialised in the reference; no particles here). */
void or_timedet_cosmo(hp_cosmo* s, double velmax, double dx)
{
    s->nstep += 1;
    double told = s->t;
    double rdt = (s->nstep > 1) ? 5.0 / (6.0 * s->dt) : 0.0;
    s->at = pow(s->t, F_TWO_THIRDS);
    s->dat = F_TWO_THIRDS * pow(s->t, F_MINUS_THIRD);
    double dta = pow((1.0 + s->ca) * s->at, 1.5) - s->t;
    s->rat = 1.0 / s->at;
    double rdtc = velmax / (s->cour * dx * s->at);
    double dt = 1.0 / dmax2(rdt, rdtc);
    dt = dmin(dt, dta);
    if (dt >= 10.0) dt = 0.1;
    double t = s->t + dt;
    if (t > s->tstop) { t = s->tstop; dt = t - told; }
    double dth = 0.5 * dt;
    double thalf = told + dth;
    s->t = t; s->dt = dt;
    s->atnew = pow(t, F_TWO_THIRDS);
    s->ath = pow(thalf, F_TWO_THIRDS);
    s->dath = F_TWO_THIRDS * pow(thalf, F_MINUS_THIRD);
    s->datnew = F_TWO_THIRDS * pow(t, F_MINUS_THIRD);
    s->rdtath = dt / s->ath;
}

/* SUBROUTINE expand (hydropad/expand.F90:1-43) over n cells, in place. */
void or_expand(double afact, double* p, double* vx, double* vy, double* vz, const double* pold,
               const double* vxold, const double* vyold, const double* vzold, long n)
{
    double afact1 = 1.0 / (1.0 + afact);
    double afact2 = 1.0 / (2.0 + afact);
    for (long i = 0; i < n; ++i) {
        vx[i] = (2.0 * vx[i] - afact * vxold[i]) * afact2;
        vy[i] = (2.0 * vy[i] - afact * vyold[i]) * afact2;
        vz[i] = (2.0 * vz[i] - afact * vzold[i]) * afact2;
        double ppp = p[i];
        p[i] = (p[i] - afact * pold[i]) * afact1;
        if (p[i] <= 0.0) p[i] = (ppp > 0.0) ? ppp : pold[i];
    }
}

/* evolve_sys with COSMO (hydropad/evolve_sys.F90:30-31,54,94-96,122-127): per
 * step keep p, vx, vy, vz, ppm_lxlylz with the schedule's scalars, expand with
 * afact = dt*(dath/ath), speed, timedet_cosmo. */
int or_evolve_cosmo(hp_params* P, double* p, double* rho, double* vx, double* vy, double* vz,
                    const double* gforce, int nx, int ny, int nz, int nb, int first_step, int nsteps,
                    hp_cosmo* s, hp_counters* cnt)
{
    int worst = 0;
    long n = (long)nx * ny * nz;
    double* old = (double*)malloc(4 * n * sizeof(double));
    if (!old) return HP_E_NOMEM;
    for (int nstep = first_step; nstep < first_step + nsteps; ++nstep) {
        int nrot = nstep % 6 + 1;
        if (nstep == 1) nrot = 1;
        P->at = s->at; P->dat = s->dat; P->rat = s->rat; P->dt = s->dt; P->rdtath = s->rdtath;
        memcpy(old, p, n * sizeof(double));
        memcpy(old + n, vx, n * sizeof(double));
        memcpy(old + 2 * n, vy, n * sizeof(double));
        memcpy(old + 3 * n, vz, n * sizeof(double));
        int r = or_ppm_lxlylz(P, p, rho, vx, vy, vz, gforce, nx, ny, nz, nb, nrot, cnt);
        if (r < 0) worst = r;
        double afact = s->dt * (s->dath / s->ath);
        or_expand(afact, p, vx, vy, vz, old, old + n, old + 2 * n, old + 3 * n, n);
        double vmax = or_speed(P, p, rho, vx, vy, vz, nx, ny, nz, nb);
        or_timedet_cosmo(s, vmax, P->dx);
    }
    P->at = s->at; P->dat = s->dat; P->rat = s->rat; P->dt = s->dt; P->rdtath = s->rdtath;
    free(old);
    return worst;
}

/* ---- N-body (nbody/kick.F90, nbody/drift.F90, hydropad/speed.F90:43-56) ---- */

/* Cell-in-cell interpolation of gforce (3,nx,ny,nz) at a particle position,
 * nbody/kick.F90:47-117 on one rank (boxlaux = 0).  ng* are the interior
 * sizes (the reference's ngridx..z).  Returns 0 when the CIC cube would leave
 * the array (the reference would read outside it). */
static int or_cic(const double* gforce, int nx, int ny, int nz, int nb, int ngx, int ngy, int ngz,
                  const double* pos, double* ff)
{
    double x1g = (pos[0] - (double)0.0f) * (double)(float)ngx;
    double x2g = (pos[1] - (double)0.0f) * (double)(float)ngy;
    double x3g = (pos[2] - (double)0.0f) * (double)(float)ngz;
    /* leading (left) cell, 1-based as the reference: int() truncates toward 0 */
    const double t1 = x1g + 0.5, t2 = x2g + 0.5, t3 = x3g + 0.5;
    if (!(t1 > -1.0 && t2 > -1.0 && t3 > -1.0 && t1 < 1e9 && t2 < 1e9 && t3 < 1e9)) return 0;
    int i1 = (int)t1 + nb, j1 = (int)t2 + nb, k1 = (int)t3 + nb;
    if (i1 < 1 || j1 < 1 || k1 < 1 || i1 + 1 > nx || j1 + 1 > ny || k1 + 1 > nz) return 0;
    double xi1 = (double)((float)(i1 - nb) - 0.5f);
    double xi2 = (double)((float)(j1 - nb) - 0.5f);
    double xi3 = (double)((float)(k1 - nb) - 0.5f);
    double dd1 = x1g - xi1, dd2 = x2g - xi2, dd3 = x3g - xi3;
    double de1 = 1.0 - dd1, de2 = 1.0 - dd2, de3 = 1.0 - dd3;
    double d1 = de1 * de2 * de3, d2 = dd1 * de2 * de3, d3 = de1 * dd2 * de3, d4 = dd1 * dd2 * de3;
    double d5 = de1 * de2 * dd3, d6 = dd1 * de2 * dd3, d7 = de1 * dd2 * dd3, d8 = dd1 * dd2 * dd3;
#define GF(c, i, j, k) gforce[(c) + 3 * ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * ((k) - 1)))]
    for (int c = 0; c < 3; ++c)
        ff[c] = GF(c, i1, j1, k1) * d1 + GF(c, i1 + 1, j1, k1) * d2 + GF(c, i1, j1 + 1, k1) * d3 +
                GF(c, i1 + 1, j1 + 1, k1) * d4 + GF(c, i1, j1, k1 + 1) * d5 + GF(c, i1 + 1, j1, k1 + 1) * d6 +
                GF(c, i1, j1 + 1, k1 + 1) * d7 + GF(c, i1 + 1, j1 + 1, k1 + 1) * d8;
#undef GF
    return 1;
}

/* SUBROUTINE kick, nbody/kick.F90:1-133: half-step velocity update
 * v <- v (1 - dat dtat/2) + F dtat/2 with F interpolated by CIC.  Particles
 * with x == -1 are inactive (:37).  Returns the number of particles skipped
 * because their CIC cube leaves the array. */
int64_t or_kick(const hp_params* P, int64_t np, const double* ppos, double* pvel, const double* gforce, int nx,
                int ny, int nz, int nb, double dtat)
{
    const double dat = P->dat;
    const int ngx = nx - 2 * nb, ngy = ny - 2 * nb, ngz = nz - 2 * nb;
    int64_t outside = 0;
    for (int64_t j = 0; j < np; ++j) {
        const double* x = ppos + 3 * j;
        double* v = pvel + 3 * j;
        if (x[0] == -1) continue;
        double ff[3];
        if (!or_cic(gforce, nx, ny, nz, nb, ngx, ngy, ngz, x, ff)) { ++outside; continue; }
        for (int c = 0; c < 3; ++c) v[c] = v[c] * (1.0 - 0.5 * dat * dtat) + 0.5 * ff[c] * dtat;
    }
    return outside;
}

/* boxsize of hydropad/init_sys.F90:93-95: ratios taken in FP32 */
static double or_boxsize(int n, int nmax) { return (double)((float)n / (float)nmax); }

/* SUBROUTINE drift, nbody/drift.F90:1-36: x <- x + rdtath v, wrapped once
 * into [0, boxsize). */
void or_drift(const hp_params* P, int64_t np, double* ppos, const double* pvel, int ngx, int ngy, int ngz)
{
    const int nmax = ngx > ngy ? (ngx > ngz ? ngx : ngz) : (ngy > ngz ? ngy : ngz);
    const double bs[3] = {or_boxsize(ngx, nmax), or_boxsize(ngy, nmax), or_boxsize(ngz, nmax)};
    for (int64_t j = 0; j < np; ++j) {
        double* x = ppos + 3 * j;
        const double* v = pvel + 3 * j;
        if (x[0] == -1) continue;
        for (int c = 0; c < 3; ++c) x[c] = x[c] + P->rdtath * v[c];
        for (int c = 0; c < 3; ++c) {
            if (x[c] < 0.0) x[c] = x[c] + bs[c];
            if (x[c] >= bs[c]) x[c] = x[c] - bs[c];
        }
    }
}

/* The particle term of SUBROUTINE speed, hydropad/speed.F90:43-56: the
 * largest |v| component of every particle (inactive ones included). */
double or_particle_vmax(int64_t np, const double* pvel)
{
    double vmax = 0.0;
    for (int64_t i = 0; i < np; ++i)
        vmax = dmax2(vmax, dmax2(dmax2(fabs(pvel[3 * i]), fabs(pvel[3 * i + 1])), fabs(pvel[3 * i + 2])));
    return vmax;
}
