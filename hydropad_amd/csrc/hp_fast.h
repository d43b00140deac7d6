// hp_fast.h -- the throughput forms of the sweep math (FP64 and mixed).
//
// Same algorithm, same branch structure and same iteration sequence as
// hp_core.h (itself the reference's operation order), with the arithmetic
// re-associated for gfx950:
//   * divisions by dx become products with rdx; repeated denominators
//     (1/C, 1/rho, 1/p of the Newton loop) are inverted once;
//   * the MC slope's m*da/|da| is copysign(m, da); the flattening ratio test
//     |da|/min >= thr is |da| >= thr*|min|; the steepening weight cancels dx;
//   * FP64 pow/log/exp are short polynomial forms (ln via atanh series,
//     exp via range-reduced Taylor) accurate to a few ulp, instead of
//     OCML's correctly-rounded-pow machinery (~224 VALU ops per call);
//   * the Newton iteration evaluates (p*/p)^m once per side and gets
//     (p*/p)^(-gf) as (p*/p)^m / (p*/p), since -gf = m - 1;
//   * the shock+rarefaction initial guess (three pow) is only formed when the
//     reference would actually use it (riemann.F90:54-56 overrides it with
//     pcheck whenever pcheck > 0).
// R = double is the FP64 form; R = float is the mixed form's reconstruction,
// tracing and Riemann sampling, with the Newton pressure iteration and the
// conservative + dual-energy update kept in FP64.
#pragma once
#include "hp_core.h"

namespace hp {
namespace fm {

// ---- primitives ------------------------------------------------------------
__device__ __forceinline__ double rcp(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(e, r, r);
    e = fma(-b, r, 1.0);
    return fma(e, r, r);
}
__device__ __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }

__device__ __forceinline__ double div(double a, double b)
{
    double r = rcp(b);
    double q = a * r;
    double e = fma(-b, q, a);
    return fma(e, r, q);
}
__device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// sqrt for x >= 0 without the range scaling of the correctly-rounded
// expansion: rsq seed and two Goldschmidt corrections (within 1 ulp for
// normal x; 0 for x == 0).
__device__ __forceinline__ double fsqrt(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return x > 0.0 ? g : 0.0;
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

constexpr double LN2_HI = 6.93147180369123816490e-01;
constexpr double LN2_LO = 1.90821492927058770002e-10;
constexpr double INV_LN2 = 1.44269504088896338700e+00;
constexpr double SQRT_HALF = 0.70710678118654752440;

// ln(x) for finite x > 0: x = 2^e f, f in [sqrt(1/2), sqrt(2)),
// ln f = 2 atanh(s), s = (f-1)/(f+1), |s| <= 0.1716, series to s^23.
__device__ __forceinline__ double ln_pos(double x, double& ehi)
{
    int e = __builtin_amdgcn_frexp_exp(x);
    double f = __builtin_amdgcn_frexp_mant(x);      // [0.5, 1)
    if (f < SQRT_HALF) { f = f + f; e -= 1; }
    double s = div(f - 1.0, f + 1.0);
    double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    p = fma(p, z, 1.0);
    ehi = (double)e;
    return 2.0 * s * p;                              // ln f; ln x = e*ln2 + this
}

// exp(y) for |y| < 700: n = rint(y/ln2), t = y - n ln2 (two-part ln2),
// Taylor to t^14 on |t| <= ln2/2.
__device__ __forceinline__ double exp_fast(double y)
{
    double n = rint(y * INV_LN2);
    double t = fma(-n, LN2_HI, y);
    t = fma(-n, LN2_LO, t);
    double p = 1.0 / 87178291200.0;                  // 1/14!
    p = fma(p, t, 1.0 / 6227020800.0);
    p = fma(p, t, 1.0 / 479001600.0);
    p = fma(p, t, 1.0 / 39916800.0);
    p = fma(p, t, 1.0 / 3628800.0);
    p = fma(p, t, 1.0 / 362880.0);
    p = fma(p, t, 1.0 / 40320.0);
    p = fma(p, t, 1.0 / 5040.0);
    p = fma(p, t, 1.0 / 720.0);
    p = fma(p, t, 1.0 / 120.0);
    p = fma(p, t, 1.0 / 24.0);
    p = fma(p, t, 1.0 / 6.0);
    p = fma(p, t, 0.5);
    p = fma(p, t, 1.0);
    p = fma(p, t, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)n);
}

// x^y for finite x > 0 (the only arguments the solver produces); 0 for x <= 0
// and y > 0.
__device__ __forceinline__ double pow_pos(double x, double y)
{
    if (!(x > 0.0)) return y > 0.0 ? 0.0 : __builtin_inf();
    double e;
    double lf = ln_pos(x, e);
    // y*ln x = y*(e*LN2_HI + (e*LN2_LO + lf)); e*LN2_HI is exact for |e| < 2^11
    double hi = e * LN2_HI;
    double lo = fma(e, LN2_LO, lf);
    return exp_fast(fma(y, hi, y * lo));
}
__device__ __forceinline__ float pow_pos(float x, float y)
{
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// ---- reconstruction ------------------------------------------------------------
template <class R> __device__ __forceinline__ R ajside(R am1, R a, R ap1)
{
    if ((ap1 - a) * (a - am1) > R(0)) {
        R da = R(0.5) * (ap1 - am1);
        R mn = tmin(tmin(fabs(da), R(2) * fabs(a - am1)), R(2) * fabs(ap1 - a));
        return copysign(mn, da);
    }
    return R(0);
}

template <class R> __device__ __forceinline__ R iface_value(R am1, R a0, R dm1, R d0)
{
    return am1 + R(0.5) * (a0 - am1) + (dm1 - d0) * R(K_SIXTH);
}

template <class R> __device__ __forceinline__ bool pair_steep(R a0, R a1, R thr)
{
    return (a1 * a0 != R(0)) && (fabs(a1 - a0) >= thr * fabs(tmin(a1, a0)));
}

// contact steepening: deltam2, deltap2 carry C6/dx^2, which cancels in eta
template <class R>
__device__ __forceinline__ void steepen(R a1, R a2, R a3, R a4, R a5, R dmalf, R dmarg, R& ajl, R& ajr)
{
    R dm = a3 - R(2) * a2 + a1;
    R dp = a5 - R(2) * a4 + a3;
    R agrad = fabs(a4 - a2) - R(K_0010) * tmin(fabs(a4), fabs(a2));
    R eta = R(0);
    if (-dm * dp > R(0) && agrad > R(0)) {
        R e = -R(K_SIXTH) * div(dp - dm, a4 - a2);
        eta = tmax(R(0), tmin(R(20) * (e - R(K_0050)), R(1)));
    }
    ajl = (R(1) - eta) * ajl + eta * (a2 + R(0.5) * dmalf);
    ajr = (R(1) - eta) * ajr + eta * (a4 - R(0.5) * dmarg);
}

// ---- characteristic tracing --------------------------------------------------------
template <class R, bool GRAV, bool COSMO>
__device__ __forceinline__ R beta_fn(const SweepScalars<R>& S, R rct, R ut, R uc, R pt, R pc, R gc, R sign)
{
    R sum = (ut - uc) + sign * (pt - pc) * rct;
    if (COSMO) sum = sum + sign * S.dt * (R(2) * S.dat * pc * rct / S.at);
    if (GRAV || COSMO) {
        R gvc = ((GRAV ? -gc : R(0)) + (COSMO ? S.dat * uc : R(0))) * S.rat;
        sum = sum + S.dt * gvc;
    }
    return -sign * sum * R(0.5) * rct;
}

template <class R, bool COSMO>
__device__ __forceinline__ R beta0_fn(const SweepScalars<R>& S, R rct, R rrhot, R rhoc, R pt, R pc)
{
    R rct2 = rct * rct;
    R b = (pt - pc) * rct2 + rrhot - rcp(rhoc);
    if (COSMO) b = b + (R(2) * S.dat * pc) * rct2 / S.at * S.dt;
    return b;
}

// right face of cell c (the reference's left state of interface c+1/2)
template <class R, bool GRAV, bool COSMO>
__device__ __forceinline__ Face<R> trace_right_face(const SweepScalars<R>& S, const Parab<R>* q, R u, R cs,
                                                    int& nclamp)
{
    const R k = S.dtrat * S.rdx;           // dt*rat/dx
    R x3 = (u + cs) * k;
    R xt = tmax(x3, R(0));
    R x1 = (u - cs) * k;
    R x2 = u * k;
    R ptl = pfl(q[VP], xt), rtl = pfl(q[VR], xt), utl = pfl(q[VU], xt);
    R vtl = pfl(q[VV], xt), wtl = pfl(q[VW], xt);
    R pc1 = pfl(q[VP], x1), pc2 = pfl(q[VP], x2), pc3 = pfl(q[VP], x3);
    R rc2 = pfl(q[VR], x2), uc1 = pfl(q[VU], x1), uc3 = pfl(q[VU], x3);
    R vc2 = pfl(q[VV], x2), wc2 = pfl(q[VW], x2);
    R gc1 = R(0), gc3 = R(0);
    if (GRAV) { gc1 = pfl(q[VG], x1); gc3 = pfl(q[VG], x3); }
    if (ptl < R(0)) { ptl = R(K_1EM15); ++nclamp; }
    if (rtl < R(0)) { rtl = R(K_1EM15); ++nclamp; }
    R ct = fsqrt(S.gamma * ptl * rtl);
    R rct = rcp(ct);
    R rrt = rcp(rtl);
    const R rat = S.rat;
    R lm = (u - cs) * rat, l0 = u * rat, lp = (u + cs) * rat;
    R bm = (lm <= R(0)) ? R(0) : fm::beta_fn<R, GRAV, COSMO>(S, rct, utl, uc1, ptl, pc1, gc1, R(-1));
    R bp = (lp <= R(0)) ? R(0) : fm::beta_fn<R, GRAV, COSMO>(S, rct, utl, uc3, ptl, pc3, gc3, R(1));
    R b0 = R(0);
    R vf = vtl, wf = wtl;
    if (l0 > R(0)) {
        b0 = fm::beta0_fn<R, COSMO>(S, rct, rrt, rc2, ptl, pc2);
        vf = vc2;
        wf = wc2;
    }
    Face<R> f;
    f.p = ptl + ct * ct * (bp + bm);
    if (f.p <= R(0)) f.p = ptl;
    f.u = utl + ct * (bp - bm);
    f.r = rcp(rrt - bm - b0 - bp);
    if (f.r <= R(0)) f.r = rtl;
    f.v = vf;
    f.w = wf;
    return f;
}

// left face of cell c (the reference's right state of interface c-1/2)
template <class R, bool GRAV, bool COSMO>
__device__ __forceinline__ Face<R> trace_left_face(const SweepScalars<R>& S, const Parab<R>* q, R u, R cs,
                                                   int& nclamp)
{
    const R k = S.dtrat * S.rdx;
    R x1 = -(u - cs) * k;
    R xt = tmax(x1, R(0));
    R x2 = -u * k;
    R x3 = -(u + cs) * k;
    R ptr = pfr(q[VP], xt), rtr = pfr(q[VR], xt), utr = pfr(q[VU], xt);
    R vtr = pfr(q[VV], xt), wtr = pfr(q[VW], xt);
    R pc1 = pfr(q[VP], x1), pc2 = pfr(q[VP], x2), pc3 = pfr(q[VP], x3);
    R rc2 = pfr(q[VR], x2), uc1 = pfr(q[VU], x1), uc3 = pfr(q[VU], x3);
    R vc2 = pfr(q[VV], x2), wc2 = pfr(q[VW], x2);
    R gc1 = R(0), gc3 = R(0);
    if (GRAV) { gc1 = pfr(q[VG], x1); gc3 = pfr(q[VG], x3); }
    if (ptr < R(0)) { ptr = R(K_1EM15); ++nclamp; }
    if (rtr < R(0)) { rtr = R(K_1EM15); ++nclamp; }
    R ct = fsqrt(S.gamma * ptr * rtr);
    R rct = rcp(ct);
    R rrt = rcp(rtr);
    const R rat = S.rat;
    R lm = (u - cs) * rat, l0 = u * rat, lp = (u + cs) * rat;
    R bp = (lp >= R(0)) ? R(0) : fm::beta_fn<R, GRAV, COSMO>(S, rct, utr, uc3, ptr, pc3, gc3, R(1));
    R bm = (lm >= R(0)) ? R(0) : fm::beta_fn<R, GRAV, COSMO>(S, rct, utr, uc1, ptr, pc1, gc1, R(-1));
    R b0 = R(0);
    R vf = vtr, wf = wtr;
    if (l0 < R(0)) {
        b0 = fm::beta0_fn<R, COSMO>(S, rct, rrt, rc2, ptr, pc2);
        vf = vc2;
        wf = wc2;
    }
    Face<R> f;
    f.p = ptr + ct * ct * (bp + bm);
    if (f.p <= R(0)) f.p = ptr;
    f.u = utr + ct * (bp - bm);
    f.r = rcp(rrt - bm - b0 - bp);
    if (f.r <= R(0)) f.r = rtr;
    f.v = vf;
    f.w = wf;
    return f;
}

// ---- Riemann solver ------------------------------------------------------------------
// Initial Newton guess of riemann.F90:54-56: pcheck when positive, else the
// caller's guess.  Returns true when pcheck was taken.
__device__ __forceinline__ bool pcheck_guess(const RScal& Rs, double pleft, double pright, double rholeft,
                                             double rhoright, double vxleft, double vxright, double& pstar)
{
    const double kappa = fsqrt(Rs.gamma * (pleft + pright) * (rholeft + rhoright) * 0.25);
    const double pcheck = 0.5 * (pleft + pright) + 0.5 * kappa * (vxleft + vxright);
    if (pcheck > 0.0) {
        pstar = pcheck;
        return true;
    }
    return false;
}

// van Leer Newton iteration in FP64 (riemann.F90:58-109) from the initial
// guess in pstar; the same iterates as the reference up to rounding.
__device__ __forceinline__ int newton(const RScal& Rs, double& pstar, double& vstar, double pleft, double pright,
                                      double rholeft, double rhoright, double vxleft, double vxright)
{
    const double gamma = Rs.gamma, m = Rs.m, gf = Rs.gf;
    const double cleft = fsqrt(gamma * pleft * rholeft);
    const double cright = fsqrt(gamma * pright * rhoright);
    const double rpl = rcp(pleft), rpr = rcp(pright);
    const double rcl = rcp(cleft), rcr = rcp(cright);
    const double cl2 = cleft * cleft, cr2 = cright * cright;
    int nnewt = 0;
    for (;;) {
        if (++nnewt > 100) return -1;
        const double pstarold = pstar;
        double a, b, vpl, vpr;
        const double xl = pstar * rpl, xr = pstar * rpr;
        if (pstar >= pleft) {
            a = cleft * fsqrt(gf * xl + m);
            vpl = -(a * a + cl2) * 0.5 * rcp(a * a * a);
        } else {
            const double w = pow_pos(xl, m);
            a = (w == 1.0) ? cleft * fsqrt(gf * xl + m) : m * cleft * div(1.0 - xl, 1.0 - w);
            vpl = -rcl * div(w, xl);
        }
        if (pstar >= pright) {
            b = cright * fsqrt(gf * xr + m);
            vpr = (b * b + cr2) * 0.5 * rcp(b * b * b);
        } else {
            const double w = pow_pos(xr, m);
            b = (w == 1.0) ? cright * fsqrt(gf * xr + m) : m * cright * div(1.0 - xr, 1.0 - w);
            vpr = rcr * div(w, xr);
        }
        const double ra = rcp(a), rb = rcp(b);
        const double vstarl = vxleft - (pstar - pleft) * ra;
        const double vstarr = vxright + (pstar - pright) * rb;
        pstar = pstar - div(vstarl - vstarr, vpl - vpr);
        if (pstar <= 0.0) pstar = pstarold * 1e-3;
        vstar = div(pleft - pright + a * vxleft + b * vxright, a + b);
        if (!(fabs(pstarold - pstar) > K_TERR * pstarold)) break;
    }
    return nnewt;
}

template <class R> __device__ __forceinline__ R rarwave(R f1, R f2, R vt, R vh)
{
    return f2 + vt * div(f1 - f2, vt - vh);
}

template <class R> __device__ __forceinline__ R rhoshock(R g1, R p2, R p1, R rho1)
{
    R x = div(p2, p1);
    return rho1 * div(g1 * x + R(1), g1 + x);
}

template <class R> __device__ __forceinline__ R vshock(R rs, R vxm, R rho, R vx)
{
    return (rs == rho) ? R(0) : div(rs * vxm - rho * vx, rs - rho);
}

// riemannsolver.F90:32-285.  Non-Newton arithmetic in R, Newton in FP64.
template <class R>
__device__ __forceinline__ Flux riemann_solve(const RScal& Rs, const Face<R>& L, const Face<R>& Rt, int& nit)
{
    const R gamma = (R)Rs.gamma, m = (R)Rs.m, rm = (R)Rs.rm, rgamma = (R)Rs.rgamma, rgamma1 = (R)Rs.rgamma1;
    const R g1 = (R)Rs.gamma1;
    const R pleft = L.p, pright = Rt.p, rholeft = L.r, rhoright = Rt.r;
    const R vxleft = L.u, vxright = Rt.u;
    R pm = R(0), rhom = R(0), vxm = R(0);
    nit = 0;
    const R cleft = fsqrt(gamma * pleft * rcp(rholeft));
    const R cright = fsqrt(gamma * pright * rcp(rhoright));
    R p1, p2, rho1, c1, c2;
    if (pleft <= pright) { p1 = pleft; p2 = pright; rho1 = rholeft; c1 = cleft; c2 = cright; }
    else { p1 = pright; p2 = pleft; rho1 = rhoright; c1 = cright; c2 = cleft; }
    const R dvv = vxleft - vxright;
    const R dv2s = div(p2 - p1, fsqrt(R(0.5) * rho1 * ((gamma - R(1)) * p1 + (gamma + R(1)) * p2)));
    const R dv2r = -R(2) * (c1 + c2) * rgamma1;
    int kase;
    if (dvv >= dv2s) kase = 1;
    else {
        const R dv1s = -(R(2) * c2 * rgamma1) * (R(1) - pow_pos(div(p1, p2), m));
        kase = (dvv >= dv1s) ? 2 : (dvv > dv2r ? 3 : 4);
    }
    if (kase == 1 || kase == 2) {
        double pmd = 0.0, vxmd = 0.0;
        if (!fm::pcheck_guess(Rs, pleft, pright, rholeft, rhoright, vxleft, vxright, pmd)) {
            if (kase == 1) pmd = (double)tmax(pleft, pright);        // riemannsolver.F90:77
            else                                                      // :117-119
                pmd = pow_pos((double)div(cleft + cright + (vxleft - vxright) * (gamma - R(1)) * R(0.5),
                                          div(cleft, pow_pos(pleft, m)) + div(cright, pow_pos(pright, m))),
                              Rs.rm);
        }
        nit = fm::newton(Rs, pmd, vxmd, pleft, pright, rholeft, rhoright, vxleft, vxright);
        pm = (R)pmd;
        vxm = (R)vxmd;
        if (kase == 1) {
            const R rs1 = rhoshock(g1, pm, pleft, rholeft);
            const R rs2 = rhoshock(g1, pm, pright, rhoright);
            const R vs1 = vshock(rs1, vxm, rholeft, vxleft);
            const R vs2 = vshock(rs2, vxm, rhoright, vxright);
            if (vs2 < R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vxm <= R(0) && vs2 >= R(0)) rhom = rs2;
            else if (vs1 <= R(0) && vxm > R(0)) rhom = rs1;
            else { pm = pleft; vxm = vxleft; rhom = rholeft; }
        } else if (pleft >= pright) {
            const R rs2 = rhoshock(g1, pm, pright, rhoright);
            const R rc1 = rholeft * pow_pos(div(pm, pleft), rgamma);
            const R vs2 = vshock(rs2, vxm, rhoright, vxright);
            const R vtail1 = vxm - fsqrt(gamma * pm * rcp(rc1));
            const R vhead1 = vxleft - cleft;
            if (vs2 < R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vhead1 >= R(0)) { pm = pleft; vxm = vxleft; rhom = rholeft; }
            else if (vxm <= R(0) && vs2 >= R(0)) rhom = rs2;
            else if (vxm > R(0) && vtail1 <= R(0)) rhom = rc1;
            else if (vtail1 > R(0) && vhead1 <= R(0)) {
                pm = fm::rarwave(pleft, pm, vtail1, vhead1);
                rhom = fm::rarwave(rholeft, rc1, vtail1, vhead1);
                vxm = fm::rarwave(vxleft, vxm, vtail1, vhead1);
            }
        } else {
            const R rs1 = rhoshock(g1, pm, pleft, rholeft);
            const R rc2 = rhoright * pow_pos(div(pm, pright), rgamma);
            const R vs1 = vshock(rs1, vxm, rholeft, vxleft);
            const R vtail2 = vxm + fsqrt(gamma * pm * rcp(rc2));
            const R vhead2 = vxright + cright;
            if (vs1 > R(0)) { pm = pleft; vxm = vxleft; rhom = rholeft; }
            else if (vhead2 <= R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vxm >= R(0) && vs1 <= R(0)) rhom = rs1;
            else if (vxm < R(0) && vtail2 >= R(0)) rhom = rc2;
            else if (vtail2 < R(0) && vhead2 >= R(0)) {
                pm = fm::rarwave(pright, pm, vtail2, vhead2);
                rhom = fm::rarwave(rhoright, rc2, vtail2, vhead2);
                vxm = fm::rarwave(vxright, vxm, vtail2, vhead2);
            }
        }
    } else if (kase == 3) {                               // two rarefactions (no Newton)
        const R zeta = cright * rcp(cleft) * pow_pos(div(pleft, pright), m);
        const R vtleft = vxleft + R(2) * cleft * rgamma1;
        const R vtright = vxright - R(2) * cright * rgamma1;
        vxm = div(vtleft * zeta + vtright, R(1) + zeta);
        pm = pleft * pow_pos(div((gamma - R(1)) * (vtleft - vtright), R(2) * cleft * (R(1) + zeta)), rm);
        const R rc1 = rholeft * pow_pos(div(pm, pleft), rgamma);
        const R rc2 = rhoright * pow_pos(div(pm, pright), rgamma);
        const R vtail2 = vxm + fsqrt(gamma * pm * rcp(rc2));
        const R vhead2 = vxright + cright;
        const R vtail1 = vxm - fsqrt(gamma * pm * rcp(rc1));
        const R vhead1 = vxleft - cleft;
        if (vhead2 <= R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vhead1 >= R(0)) { pm = pleft; vxm = vxleft; rhom = rholeft; }
        else if (vhead2 > R(0) && vtail2 <= R(0)) {
            pm = fm::rarwave(pright, pm, vtail2, vhead2);
            rhom = fm::rarwave(rhoright, rc2, vtail2, vhead2);
            vxm = fm::rarwave(vxright, vxm, vtail2, vhead2);
        } else if (vxm <= R(0) && vtail2 > R(0)) rhom = rc2;
        else if (vxm > R(0) && vtail1 <= R(0)) rhom = rc1;
        else if (vtail1 > R(0) && vhead1 <= R(0)) {
            pm = fm::rarwave(pleft, pm, vtail1, vhead1);
            rhom = fm::rarwave(rholeft, rc1, vtail1, vhead1);
            vxm = fm::rarwave(vxleft, vxm, vtail1, vhead1);
        }
    } else {                                              // vacuum
        pm = R(0);
        vxm = R(0);
        const R vtail2 = vxright - R(2) * cright * rgamma1;
        const R vhead2 = vxright + cright;
        const R vtail1 = vxleft - R(2) * cleft * rgamma1;
        const R vhead1 = vxleft - cleft;
        if (vhead2 <= R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vhead1 >= R(0)) { pm = pleft; vxm = vxleft; rhom = rholeft; }
        else if (vhead2 > R(0) && vtail2 <= R(0)) {
            pm = fm::rarwave(pright, pm, vtail2, vhead2);
            rhom = fm::rarwave(rhoright, R(0), vtail2, vhead2);
            vxm = fm::rarwave(vxright, vxm, vtail2, vhead2);
        } else if (vtail1 <= R(0) && vtail2 > R(0)) rhom = R(0);
        else if (vtail1 > R(0) && vhead1 <= R(0)) {
            pm = fm::rarwave(pleft, pm, vtail1, vhead1);
            rhom = fm::rarwave(rholeft, R(0), vtail1, vhead1);
            vxm = fm::rarwave(vxleft, vxm, vtail1, vhead1);
        }
    }
    Flux F;
    F.p = (double)pm;
    F.r = (double)rhom;
    F.u = (double)vxm;
    if (vxm >= R(0)) { F.v = Rs.dadt * (double)L.v; F.w = Rs.dadt * (double)L.w; }
    else { F.v = Rs.dadt * (double)Rt.v; F.w = Rs.dadt * (double)Rt.w; }
    return F;
}

// ---- conservative update + dual energy (FP64) ------------------------------------------
__device__ __forceinline__ void integrate_cell(const IScal& S, const Flux& F1, const Flux& F2, double& pres,
                                               double& rho, double& vx, double& vy, double& vz, double etot,
                                               double ghalf, bool grav)
{
    const double lam = S.rdtath * S.rdx, gamma = S.gamma, rgamma1 = S.rgamma1;
    const double eint = pres * rgamma1;
    const double m1 = F1.r * F1.u, m2 = F2.r * F2.u;     // mass fluxes
    double rhonew = rho + lam * (m1 - m2);
    const double rrn = rcp(rhonew);
    double px = vx * rho + lam * ((m1 * F1.u + F1.p) - (m2 * F2.u + F2.p));
    if (grav) px += S.rdtath * ((rho + rhonew) * ghalf * 0.5);
    const double vxnew = px * rrn;
    const double vynew = (rho * vy + lam * (F1.v * m1 - F2.v * m2)) * rrn;
    const double vznew = (rho * vz + lam * (F1.w * m1 - F2.w * m2)) * rrn;
    const double e1 = (gamma * F1.p * rgamma1 + 0.5 * F1.r * (F1.u * F1.u + F1.v * F1.v + F1.w * F1.w)) * F1.u;
    const double e2 = (gamma * F2.p * rgamma1 + 0.5 * F2.r * (F2.u * F2.u + F2.v * F2.v + F2.w * F2.w)) * F2.u;
    double etotnew = etot + lam * (e1 - e2);
    if (grav) etotnew += S.rdtath * ((rho * vx + rhonew * vxnew) * ghalf * 0.5);
    const double pmean = (F1.p + F2.p) * 0.5;
    double eintnew = eint + lam * (pmean * (F1.u - F2.u) + (F1.p * F1.u - F2.p * F2.u) * rgamma1);
    if (eintnew < 0.0) eintnew = eint;
    double pnew;
    dual_energy(S, etotnew, rhonew, vxnew, vynew, vznew, eint, pres, eintnew, pnew);
    if (rhonew <= 0.0) rhonew = rho;
    pres = pnew;
    rho = rhonew;
    vx = vxnew;
    vy = vynew;
    vz = vznew;
}

}  // namespace fm
}  // namespace hp
