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
//     (p*/p)^(-gf) = (p*/p)^(m-1) from the same root (for gamma = 5/3 the
//     fifth root's own y^4); its step quotient skips the residual
//     correction, which only scales the last step;
//   * the shock+rarefaction initial guess (three pow) is only formed when the
//     reference would actually use it (riemann.F90:54-56 overrides it with
//     pcheck whenever pcheck > 0);
//   * a contact at rest (equal pressures, no normal velocity) is returned
//     exactly without iterating (the value the full path gives).
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

// a/b: one Newton step on v_rcp_f64 (accurate to about 2^-24 on gfx950, so
// 2^-48 after the step), then one residual correction of the quotient, which
// squares that error again -- within an ulp without rcp's second step.
__device__ __forceinline__ double div(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    const double e = fma(-b, q, a);
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

// 1/sqrt(x) and sqrt(x) for x > 0 together: v_rsq_f64 seed and two Newton
// steps on 1/sqrt, y <- y (1.5 - 0.5 x y^2).
__device__ __forceinline__ void rsq_sqrt(double x, double& rs, double& sq)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * (x * y), y, 1.5);
    y = y * fma(-0.5 * (x * y), y, 1.5);
    rs = y;
    sq = x * y;
}

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

// x^(1/5) for x >= 0: FP32 seed of y = x^(-1/5), two Newton steps on
// y^-5 = x, y <- y (6 - x y^5)/5 (no division; each step takes the relative
// error e to about 3e^2), then x^(1/5) = x y^4.  The seed needs x inside
// FP32's normal range; near-vacuum rarefactions produce pressure ratios far
// below it (p*/p ~ 1e-45 and less), so outside [1e-30, 1e30] the argument is
// first scaled by 2^(-5q) and the root by 2^q.  x = 0 gives 0.  rx45
// receives x^(-4/5) = x^(1/5)/x, the y^4 on the way (0 for x = 0).
// The common case runs unconditionally and the rare one overwrites its
// results, so the in-range path carries no merge moves.
__device__ __forceinline__ double root5_core(double xs, double& y4)
{
    double y = (double)__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf((float)xs));
    double y2 = y * y;
    y = y * fma(-0.2 * xs, y2 * y2 * y, 1.2);
    y2 = y * y;
    y = y * fma(-0.2 * xs, y2 * y2 * y, 1.2);
    y2 = y * y;
    y4 = y2 * y2;
    return xs * y4;
}
__device__ __forceinline__ double root5(double x, double& rx45)
{
    double r = root5_core(x, rx45);
    if (!(x > 1e-30 && x < 1e30)) {                 // rare: wave-uniformly skipped
        if (!(x > 0.0)) {
            rx45 = 0.0;
            r = 0.0;
        } else {
            const int e = __builtin_amdgcn_frexp_exp(x);
            const int q = (e >= 0) ? e / 5 : -((4 - e) / 5);
            double y4;
            r = root5_core(__builtin_amdgcn_ldexp(x, -5 * q), y4) * __builtin_amdgcn_ldexp(1.0, q);
            rx45 = y4 * __builtin_amdgcn_ldexp(1.0, -4 * q);
        }
    }
    return r;
}
__device__ __forceinline__ double root5(double x)
{
    double rx45;
    return root5(x, rx45);
}

// The three powers the Riemann solver takes: x^m, x^(1/gamma), x^(1/m).  For
// gamma = 5/3 (m = 1/5) they are x^(1/5), (x^(1/5))^3 and x^5.
__device__ __forceinline__ double pow_m(const RScal& R, double x) { return R.m5 ? root5(x) : pow_pos(x, R.m); }
__device__ __forceinline__ double pow_rgamma(const RScal& R, double x)
{
    if (R.m5) {
        const double y = root5(x);
        return y * y * y;
    }
    return pow_pos(x, R.rgamma);
}
__device__ __forceinline__ double pow_rm(const RScal& R, double x)
{
    if (R.m5) {
        const double x2 = x * x;
        return x2 * x2 * x;
    }
    return pow_pos(x, R.rm);
}

// ---- reconstruction and tracing, delta form ---------------------------------------
// Cell values stay FP64.  R (float in the mixed form, double in the FP64 form)
// carries only differences and offsets from a cell's own value: one-sided
// differences, MC slopes, the parabola's edge offsets (ajl - a, ajr - a) and
// the characteristic corrections.  The parabola algebra is translation
// invariant (a6 = 6(a - (ajl+ajr)/2) = -3(l+r) in offsets; the CW84 limits
// become l = -2r / r = -2l; flattening scales the offsets by 1-flat), so FP32
// rounding acts on the small offsets, not on the O(1) values.  Every switch
// whose flip would be a jump -- flattening, the steepening switch, the
// characteristic signs -- is decided in FP64 from the FP64 values.

// MC slope (ppm_functs.F90:167-181) from the cell's one-sided differences.
template <class R> __device__ __forceinline__ R slope_d(double am1, double a, double ap1)
{
    const R dm = (R)(a - am1), dp = (R)(ap1 - a);
    if (dp * dm > R(0)) {
        const R da = R(0.5) * (dm + dp);
        const R mn = tmin(tmin(fabs(da), R(2) * fabs(dm)), R(2) * fabs(dp));
        return copysign(mn, da);
    }
    return R(0);
}

// interface value between (c-1 | c) before limiting, FP64 (ppm_functs.F90:33-34)
template <class R> __device__ __forceinline__ double iface_d(double am1, double a0, R dm1, R d0)
{
    return am1 + 0.5 * (a0 - am1) + (double)((dm1 - d0) * R(K_SIXTH));
}

// one pair of the flattening test, FP64 (order2interp.F90:65-113)
__device__ __forceinline__ bool pair_steep_d(double a0, double a1, double thr)
{
    return (a1 * a0 != 0.0) && (fabs(a1 - a0) >= thr * fabs(tmin(a1, a0)));
}

// contact steepening on offsets (dda.F90:182-216, ppm_functs.F90:35-40);
// a1..a5 = rho(c-2..c+2), dmalf/dmarg the slopes of c-1 / c+1
template <class R>
__device__ __forceinline__ void steepen_d(double a1, double a2, double a3, double a4, double a5, R dmalf, R dmarg,
                                          R& l, R& r)
{
    const double dm = (a3 - a2) - (a2 - a1);          // second differences; C6/dx^2 cancels in eta
    const double dp = (a5 - a4) - (a4 - a3);
    const double agrad = fabs(a4 - a2) - K_0010 * tmin(fabs(a4), fabs(a2));
    R eta = R(0);
    if (-dm * dp > 0.0 && agrad > 0.0) {
        const R e = -R(K_SIXTH) * div((R)(dp - dm), (R)(a4 - a2));
        eta = tmax(R(0), tmin(R(20) * (e - R(K_0050)), R(1)));
    }
    l = (R(1) - eta) * l + eta * ((R)(a2 - a3) + R(0.5) * dmalf);
    r = (R(1) - eta) * r + eta * ((R)(a4 - a3) - R(0.5) * dmarg);
}

// CW84 monotonicity + flattening on edge offsets (ppm_functs.F90:41-54)
template <class R> __device__ __forceinline__ Parab<R> limit_parab_d(R l, R r, R flat)
{
    R d = r - l;
    R a6 = R(-3) * (l + r);
    if (-(r * l) <= R(0)) {
        l = R(0);
        r = R(0);
    } else if (d * a6 > d * d) {
        l = R(-2) * r;
    } else if (d * a6 < (-(d * d))) {
        r = R(-2) * l;
    }
    l = (R(1) - flat) * l;
    r = (R(1) - flat) * r;
    Parab<R> q;
    q.l = l;
    q.r = r;
    q.d = r - l;
    q.a6 = R(-3) * (l + r);
    return q;
}

// beta (ppm_functs.F90:129-144) on offsets: ut-uc and pt-pc are offset
// differences; pc/uc/gc absolute values enter only the expansion/gravity terms.
template <class R, bool GRAV, bool COSMO>
__device__ __forceinline__ R beta_d(const SweepScalars<R>& S, R rct, R ut, R uc, R pt, R pc, R pc_abs, R uc_abs,
                                    R gc_abs, R sign)
{
    R sum = (ut - uc) + sign * (pt - pc) * rct;
    if (COSMO) sum = sum + sign * (S.cdat2 * pc_abs * rct);            // dt 2 adot pc/(C a)
    if (GRAV || COSMO) {
        R gvc = ((GRAV ? -gc_abs : R(0)) + (COSMO ? S.dat * uc_abs : R(0))) * S.rat;
        sum = sum + S.dt * gvc;
    }
    return -sign * sum * R(0.5) * rct;
}

// Traced face state of a cell.  RIGHT = true: its right face (the
// reference's left state of interface c+1/2, fl of the parabola at the
// characteristics u+c, u, u-c, ppm_functs.F90:59-75, order2interp.F90:192-219);
// RIGHT = false: its left face (the right state of c-1/2, fr at -(u-c), -u,
// -(u+c), ppm_functs.F90:62-84, order2interp.F90:221-247).
template <class R, bool GRAV, bool COSMO, bool RIGHT>
__device__ __forceinline__ Face<double> trace_face_d(const SweepScalars<R>& S, const Parab<R>* q, const double* a,
                                                     double cs, double gamma, int& nclamp)
{
    const double um = a[VU] - cs, u0 = a[VU], up = a[VU] + cs;     // characteristic speeds (FP64 signs)
    const R k = S.dtrat * S.rdx;
    R xt, x1, x2, x3;
    if (RIGHT) {
        x3 = (R)up * k;
        xt = tmax(x3, R(0));
        x1 = (R)um * k;
        x2 = (R)u0 * k;
    } else {
        x1 = -(R)um * k;
        xt = tmax(x1, R(0));
        x2 = -(R)u0 * k;
        x3 = -(R)up * k;
    }
    auto tr = [&](const Parab<R>& p, R x) { return RIGHT ? pfl(p, x) : pfr(p, x); };
    R pt = tr(q[VP], xt), rt = tr(q[VR], xt), ut = tr(q[VU], xt), vt = tr(q[VV], xt), wt = tr(q[VW], xt);
    const R pc1 = tr(q[VP], x1), pc2 = tr(q[VP], x2), pc3 = tr(q[VP], x3);
    const R rc2 = tr(q[VR], x2), uc1 = tr(q[VU], x1), uc3 = tr(q[VU], x3);
    const R vc2 = tr(q[VV], x2), wc2 = tr(q[VW], x2);
    R gc1 = R(0), gc3 = R(0);
    if (GRAV) {
        gc1 = (R)a[VG] + tr(q[VG], x1);
        gc3 = (R)a[VG] + tr(q[VG], x3);
    }
    double ptl = a[VP] + (double)pt, rtl = a[VR] + (double)rt;
    if (ptl < 0.0) { ptl = K_1EM15; pt = (R)(K_1EM15 - a[VP]); ++nclamp; }
    if (rtl < 0.0) { rtl = K_1EM15; rt = (R)(K_1EM15 - a[VR]); ++nclamp; }
    // the Lagrangian sound speed and its inverse (order2interp.F90:192-193):
    // in FP64 both from one rsq; in FP32 the hardware root and reciprocal
    R ct, rct;
    if constexpr (sizeof(R) == 8) {
        rsq_sqrt(gamma * ptl * rtl, rct, ct);
    } else {
        ct = fsqrt((R)(gamma * ptl * rtl));
        rct = rcp(ct);
    }
    const R ap = (R)a[VP], au = (R)a[VU];
    R bm = R(0), bp = R(0), b0 = R(0);
    R vf = vt, wf = wt;
    if (RIGHT) {
        if (um > 0.0) bm = beta_d<R, GRAV, COSMO>(S, rct, ut, uc1, pt, pc1, ap + pc1, au + uc1, gc1, R(-1));
        if (up > 0.0) bp = beta_d<R, GRAV, COSMO>(S, rct, ut, uc3, pt, pc3, ap + pc3, au + uc3, gc3, R(1));
        if (u0 > 0.0) {
            vf = vc2;
            wf = wc2;
        }
    } else {
        if (up < 0.0) bp = beta_d<R, GRAV, COSMO>(S, rct, ut, uc3, pt, pc3, ap + pc3, au + uc3, gc3, R(1));
        if (um < 0.0) bm = beta_d<R, GRAV, COSMO>(S, rct, ut, uc1, pt, pc1, ap + pc1, au + uc1, gc1, R(-1));
        if (u0 < 0.0) {
            vf = vc2;
            wf = wc2;
        }
    }
    if (RIGHT ? (u0 > 0.0) : (u0 < 0.0)) {
        // beta0 (ppm_functs.F90:149-162) with 1/rho_t - 1/rho_c = (rho_c - rho_t)/(rho_t rho_c)
        const R rct2 = rct * rct;
        const R rtr = (R)rtl, rcr = (R)a[VR] + rc2;
        b0 = (pt - pc2) * rct2 + div(rc2 - rt, rtr * rcr);
        if (COSMO) b0 = b0 + S.cdat2 * (ap + pc2) * rct2;                 // 2 adot pc dt/(a C^2)
    }
    Face<double> f;
    f.p = ptl + (double)(ct * ct * (bp + bm));
    if (f.p <= 0.0) f.p = ptl;
    f.u = a[VU] + (double)(ut + ct * (bp - bm));
    // 1/(1/rho_t - bm - b0 - bp) = rho_t / (1 - rho_t (bm + b0 + bp))
    if constexpr (sizeof(R) == 8) {
        f.r = rtl * rcp(1.0 - (double)((R)rtl * (bm + b0 + bp)));
    } else {                                        // delta form: rtl + rtl s/(1 - s)
        const R sb = (R)rtl * (bm + b0 + bp);
        f.r = rtl + rtl * (double)(sb * rcp(R(1) - sb));
    }
    if (f.r <= 0.0) f.r = rtl;
    f.v = a[VV] + (double)vf;
    f.w = a[VW] + (double)wf;
    return f;
}

// Parabolas of one cell and its two traced faces (delta form).
//   a[]          the cell's FP64 values (p, rho, u, v, w[, g])
//   ivl[], ivr[] FP64 interface values before limiting at its left / right interface
//   rm2..rp2     rho of cells c-2..c+2; drm1/drp1 rho slopes of c-1, c+1
//   pm1, pp1     p of cells c-1, c+1
template <class R, int NV, bool GRAV, bool COSMO>
__device__ __forceinline__ void cell_faces_d(const SweepScalars<R>& S, double gamma, const double* a,
                                             const double* ivl, const double* ivr, double rm2, double rm1,
                                             double rp1, double rp2, R drm1, R drp1, double pm1, double pp1,
                                             bool flat, Face<double>& fL, Face<double>& fR, int& nclL, int& nclR)
{
    const R flatv = flat ? S.flatvalue : R(0);
    // steepening switch addrho >= addp (order2interp.F90:152-153, ppm_functs.F90:35),
    // FP64 and cross-multiplied: both minima are positive
    const bool steep_rho = (0.5 * gamma * fabs(rp1 - rm1)) * tmin(pm1, pp1) >= fabs(pp1 - pm1) * tmin(rm1, rp1);
    Parab<R> q[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        R l = (R)(ivl[v] - a[v]), r = (R)(ivr[v] - a[v]);
        if (v == VR && steep_rho) steepen_d(rm2, rm1, a[VR], rp1, rp2, drm1, drp1, l, r);
        q[v] = limit_parab_d(l, r, flatv);
    }
    // ppmsolver.F90:109; the mixed form takes it at single-precision accuracy
    // (raw v_rcp_f64 / v_rsq_f64): it sets the characteristic speeds u -+ c,
    // whose signs only switch corrections that vanish where u -+ c does
    double cs;
    if constexpr (sizeof(R) == 8) {
        cs = fsqrt(gamma * a[VP] * rcp(a[VR]));
    } else {
        const double c2 = gamma * a[VP] * __builtin_amdgcn_rcp(a[VR]);
        cs = c2 > 0.0 ? c2 * __builtin_amdgcn_rsq(c2) : 0.0;
    }
    fR = trace_face_d<R, GRAV, COSMO, true>(S, q, a, cs, gamma, nclR);
    fL = trace_face_d<R, GRAV, COSMO, false>(S, q, a, cs, gamma, nclL);
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


// One side of the van Leer iteration at pstar (riemann.F90:70-97): the inverse
// wave impedance 1/a and the left side's form of dv*/dp* (the right side's is
// its negative).  shock = pstar >= p_side; cl2 = c^2, rc = 1/c for the side's
// Lagrangian sound speed c; x = pstar/p_side, xinv = p_side/pstar.
__device__ __forceinline__ void newton_side(const RScal& Rs, bool shock, double x, double pstar, double pside,
                                            double cl2, double rc, double& ra, double& vp)
{
    const double m = Rs.m, gf = Rs.gf;
    double w = 1.0, wx = 0.0;                        // x^m and x^m / x
    if (!shock) {
        if (Rs.m5) {
            w = root5(x, wx);
        } else {
            w = pow_pos(x, m);
            wx = w * (pside * rcp(pstar));
        }
        shock = (w == 1.0);                          // pstup == 0: the shock impedance (:74-75)
    }
    if (shock) {                                     // a = c sqrt(gf x + m), v' = -(a^2 + c^2)/(2 a^3)
        const double q = fma(gf, x, m);
        double rs, sq;
        rsq_sqrt(q, rs, sq);
        ra = rc * rs;
        vp = -(cl2 * q + cl2) * 0.5 * (ra * ra * ra);
    } else {                                         // a = m c (1 - x)/(1 - x^m), v' = -(1/c) x^m / x
        ra = (1.0 - w) * rcp(m * (1.0 - x)) * rc;
        vp = -rc * wx;
    }
}

// van Leer Newton iteration in FP64 (riemann.F90:58-109) from the initial
// iterate in pstar; the same iterates as the reference up to rounding.  Runs
// until the relative change is within terr and returns the iterations done,
// or -1 past the reference's limit (where riemann.F90:60-65 stops).  vstar
// is the reference's (pl - pr + a vl + b vr)/(a + b) with the last
// iteration's impedances, written with 1/a and 1/b.  cleft/cright are the
// Lagrangian sound speeds sqrt(gamma p rho) of the two sides
// (riemann.F90:68-69), rcl/rcr their inverses.
// The loop's limit is the reference's constant nnewtmax = 100; a different
// limit set through the API enters as a head start of the counter
// (Rs.nskip = 100 - limit), which keeps a runtime bound out of the loop.
__device__ __forceinline__ int newton(const RScal& Rs, double& pstar, double& vstar, double pleft, double pright,
                                      double vxleft, double vxright, double cleft, double rcl, double cright,
                                      double rcr)
{
    const double cl2 = cleft * cleft, cr2 = cright * cright;
    const double rpl = rcp(pleft), rpr = rcp(pright);
    double ra = 0.0, rb = 0.0;
    int k = Rs.nskip;
    for (;;) {
        if (++k > K_NNEWTMAX) {                      // where the reference stops (riemann.F90:60-65)
            k = -1;
            break;
        }
        const double pstarold = pstar;
        double vpl, vpr;
        newton_side(Rs, pstar >= pleft, pstar * rpl, pstar, pleft, cl2, rcl, ra, vpl);
        newton_side(Rs, pstar >= pright, pstar * rpr, pstar, pright, cr2, rcr, rb, vpr);
        const double vstarl = vxleft - (pstar - pleft) * ra;
        const double vstarr = vxright + (pstar - pright) * rb;
        // the step (vprimel - vprimer = vpl + vpr) with one Newton step on the
        // reciprocal and no quotient correction: within 2^-47 of the step,
        // which moves the final iterate by at most that much of its last step
        const double d = vpl + vpr;
        double rd = __builtin_amdgcn_rcp(d);
        rd = fma(fma(-d, rd, 1.0), rd, rd);
        pstar = pstar - (vstarl - vstarr) * rd;
        if (pstar <= 0.0) pstar = pstarold * 1e-3;
        if (!(fabs(pstarold - pstar) > K_TERR * pstarold)) break;
    }
    vstar = div(fma((pleft - pright) * ra, rb, fma(vxleft, rb, vxright * ra)), ra + rb);
    return k < 0 ? -1 : k - Rs.nskip;
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

// A Riemann problem after the wave-pattern classification of
// riemannsolver.F90:32-64: Eulerian sound speeds and inverses, the pattern,
// and for patterns 1 and 2 the Newton iterate (initial guess, then result).
struct RSolve {
    double rrl, rrr, rcle, cleft, rcre, cright;
    int kase, nit;
    double pm, vxm;
};

// Classification and initial Newton guess (riemannsolver.F90:32-64,77,117-119,
// riemann.F90:54-56).
__device__ __forceinline__ void riemann_setup(const RScal& Rs, const Face<double>& L, const Face<double>& Rt,
                                              RSolve& S)
{
    using R = double;
    const R gamma = (R)Rs.gamma, rgamma1 = (R)Rs.rgamma1;
    const R pleft = L.p, pright = Rt.p, rholeft = L.r, rhoright = Rt.r;
    const R vxleft = L.u, vxright = Rt.u;
    // Eulerian sound speeds and their inverses from one rsq each; the
    // Newton iteration's Lagrangian C = rho c and 1/C = (1/c)(1/rho) follow
    S.rrl = rcp(rholeft);
    S.rrr = rcp(rhoright);
    rsq_sqrt(gamma * pleft * S.rrl, S.rcle, S.cleft);
    rsq_sqrt(gamma * pright * S.rrr, S.rcre, S.cright);
    const double cleft = S.cleft, cright = S.cright;
    R p1, p2, rho1, c1, c2;
    if (pleft <= pright) { p1 = pleft; p2 = pright; rho1 = rholeft; c1 = cleft; c2 = cright; }
    else { p1 = pright; p2 = pleft; rho1 = rhoright; c1 = cright; c2 = cleft; }
    const R dvv = vxleft - vxright;
    double rs2, sq2;
    rsq_sqrt(R(0.5) * rho1 * ((gamma - R(1)) * p1 + (gamma + R(1)) * p2), rs2, sq2);
    const R dv2s = (p2 - p1) * rs2;
    const R dv2r = -R(2) * (c1 + c2) * rgamma1;
    int kase;
    if (dvv >= dv2s) kase = 1;
    else {
        const R dv1s = -(R(2) * c2 * rgamma1) * (R(1) - pow_m(Rs, div(p1, p2)));
        kase = (dvv >= dv1s) ? 2 : (dvv > dv2r ? 3 : 4);
    }
    S.kase = kase;
    S.nit = 0;
    S.pm = 0.0;
    S.vxm = 0.0;
    if (kase == 1 || kase == 2) {
        double pmd = 0.0;
        if (!fm::pcheck_guess(Rs, pleft, pright, rholeft, rhoright, vxleft, vxright, pmd)) {
            if (kase == 1) pmd = (double)tmax(pleft, pright);        // riemannsolver.F90:77
            else                                                      // :117-119
                pmd = pow_rm(Rs, div(cleft + cright + (vxleft - vxright) * (gamma - R(1)) * R(0.5),
                                     div(cleft, pow_m(Rs, pleft)) + div(cright, pow_m(Rs, pright))));
        }
        S.pm = pmd;
    }
}

// Newton iterations of patterns 1 and 2 from S.pm: S.pm, S.vxm the star
// state, S.nit the iterations (-1 where the reference stops).
__device__ __forceinline__ void riemann_newton(const RScal& Rs, const Face<double>& L, const Face<double>& Rt,
                                               RSolve& S)
{
    S.nit = newton(Rs, S.pm, S.vxm, L.p, Rt.p, L.u, Rt.u, L.r * S.cleft, S.rcle * S.rrl, Rt.r * S.cright,
                   S.rcre * S.rrr);
}

// The costly sampling operations in precision Q.  Q = double: the FP64
// form's (the reference's algebra, divisions and roots to within an ulp).
// Q = float: the mixed form's, at single-precision accuracy -- the raw
// v_rcp_f64 / v_rsq_f64 results (about 2^-24 relative) without the Newton
// refinement, in FP64 registers, so no value leaves FP64's range -- for the
// shock density, the star-region sound speed and the fan interpolation
// weight.  Each of those scales a difference (the density jump across the
// shock, the distance of a wave speed from 0, f1 - f2) or only steers a
// switch, so its rounding stays at single-precision relative size of that
// difference.  The rarefaction density rho (pm/p)^(1/gamma) and the
// two-rarefaction star state stay at FP64 accuracy: single-precision powers
// there round the full state, and over 300 steps of the 512^3 Sod tube that
// raised the mixed form's L1(rho) distance to the FP64 form from 1e-7 to 2e-5.
template <class Q> struct SampleOps;
template <> struct SampleOps<double> {
    static __device__ __forceinline__ double shock_density(double g1, double pm, double p, double rho)
    {
        return rhoshock(g1, pm, p, rho);
    }
    static __device__ __forceinline__ double fan_density(const RScal& Rs, double pm, double p, double rho)
    {
        return rho * pow_rgamma(Rs, div(pm, p));
    }
    static __device__ __forceinline__ double sound(double gamma, double p, double rho)
    {
        return fsqrt(gamma * p * rcp(rho));
    }
    static __device__ __forceinline__ double fan(double f1, double f2, double vt, double vh)
    {
        return fm::rarwave(f1, f2, vt, vh);
    }
    static __device__ __forceinline__ double ratio(double a, double b) { return div(a, b); }
    static __device__ __forceinline__ double powm_ratio(const RScal& Rs, double a, double b)
    {
        return pow_m(Rs, div(a, b));
    }
};
template <> struct SampleOps<float> : SampleOps<double> {
    // Rankine-Hugoniot (riemann.F90:22) in delta form: rho (g1 x + 1)/(g1 + x)
    // = rho + rho (g1 - 1)(pm - p)/(g1 p + pm), the approximate quotient
    // scaling the jump only (exactly rho when pm = p)
    static __device__ __forceinline__ double shock_density(double g1, double pm, double p, double rho)
    {
        return rho + rho * ((g1 - 1.0) * ((pm - p) * __builtin_amdgcn_rcp(g1 * p + pm)));
    }
    static __device__ __forceinline__ double sound(double gamma, double p, double rho)
    {
        const double c2 = gamma * p * __builtin_amdgcn_rcp(rho);
        return c2 > 0.0 ? c2 * __builtin_amdgcn_rsq(c2) : 0.0;
    }
    // riemann.F90:7: f2 + (f1 - f2) vt/(vt - vh), the weight approximate
    static __device__ __forceinline__ double fan(double f1, double f2, double vt, double vh)
    {
        return f2 + (vt * __builtin_amdgcn_rcp(vt - vh)) * (f1 - f2);
    }
};

// Sampling at x/t = 0 (riemannsolver.F90:73-275) from the classified problem
// and, for patterns 1 and 2, the Newton result; the costly operations in
// precision Q (SampleOps), switch decisions in FP64.
template <class Q>
__device__ __forceinline__ Flux riemann_sample(const RScal& Rs, const Face<double>& L, const Face<double>& Rt,
                                               const RSolve& S)
{
    using R = double;
    using O = SampleOps<Q>;
    const R gamma = (R)Rs.gamma, m = (R)Rs.m, rm = (R)Rs.rm, rgamma = (R)Rs.rgamma, rgamma1 = (R)Rs.rgamma1;
    const R g1 = (R)Rs.gamma1;
    (void)m; (void)rm; (void)rgamma;
    const R pleft = L.p, pright = Rt.p, rholeft = L.r, rhoright = Rt.r;
    const R vxleft = L.u, vxright = Rt.u;
    const double cleft = S.cleft, cright = S.cright, rcle = S.rcle;
    const int kase = S.kase;
    R pm = R(0), rhom = R(0), vxm = R(0);
    if (kase == 0) {                                      // contact at rest (riemann_solve)
        pm = pleft;
        rhom = rhoright;
    } else if (kase == 1 || kase == 2) {
        pm = (R)S.pm;
        vxm = (R)S.vxm;
        if (kase == 1) {
            const R rs1 = O::shock_density(g1, pm, pleft, rholeft);
            const R rs2 = O::shock_density(g1, pm, pright, rhoright);
            const R vs1 = vshock(rs1, vxm, rholeft, vxleft);
            const R vs2 = vshock(rs2, vxm, rhoright, vxright);
            if (vs2 < R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vxm <= R(0) && vs2 >= R(0)) rhom = rs2;
            else if (vs1 <= R(0) && vxm > R(0)) rhom = rs1;
            else { pm = pleft; vxm = vxleft; rhom = rholeft; }
        } else if (pleft >= pright) {
            const R rs2 = O::shock_density(g1, pm, pright, rhoright);
            const R rc1 = O::fan_density(Rs, pm, pleft, rholeft);
            const R vs2 = vshock(rs2, vxm, rhoright, vxright);
            const R vtail1 = vxm - O::sound(gamma, pm, rc1);
            const R vhead1 = vxleft - cleft;
            if (vs2 < R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vhead1 >= R(0)) { pm = pleft; vxm = vxleft; rhom = rholeft; }
            else if (vxm <= R(0) && vs2 >= R(0)) rhom = rs2;
            else if (vxm > R(0) && vtail1 <= R(0)) rhom = rc1;
            else if (vtail1 > R(0) && vhead1 <= R(0)) {
                pm = O::fan(pleft, pm, vtail1, vhead1);
                rhom = O::fan(rholeft, rc1, vtail1, vhead1);
                vxm = O::fan(vxleft, vxm, vtail1, vhead1);
            }
        } else {
            const R rs1 = O::shock_density(g1, pm, pleft, rholeft);
            const R rc2 = O::fan_density(Rs, pm, pright, rhoright);
            const R vs1 = vshock(rs1, vxm, rholeft, vxleft);
            const R vtail2 = vxm + O::sound(gamma, pm, rc2);
            const R vhead2 = vxright + cright;
            if (vs1 > R(0)) { pm = pleft; vxm = vxleft; rhom = rholeft; }
            else if (vhead2 <= R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
            else if (vxm >= R(0) && vs1 <= R(0)) rhom = rs1;
            else if (vxm < R(0) && vtail2 >= R(0)) rhom = rc2;
            else if (vtail2 < R(0) && vhead2 >= R(0)) {
                pm = O::fan(pright, pm, vtail2, vhead2);
                rhom = O::fan(rhoright, rc2, vtail2, vhead2);
                vxm = O::fan(vxright, vxm, vtail2, vhead2);
            }
        }
    } else if (kase == 3) {                               // two rarefactions (no Newton)
        const R zeta = cright * rcle * O::powm_ratio(Rs, pleft, pright);
        const R vtleft = vxleft + R(2) * cleft * rgamma1;
        const R vtright = vxright - R(2) * cright * rgamma1;
        vxm = O::ratio(vtleft * zeta + vtright, R(1) + zeta);
        pm = pleft * pow_rm(Rs, O::ratio((gamma - R(1)) * (vtleft - vtright), R(2) * cleft * (R(1) + zeta)));
        const R rc1 = O::fan_density(Rs, pm, pleft, rholeft);
        const R rc2 = O::fan_density(Rs, pm, pright, rhoright);
        const R vtail2 = vxm + O::sound(gamma, pm, rc2);
        const R vhead2 = vxright + cright;
        const R vtail1 = vxm - O::sound(gamma, pm, rc1);
        const R vhead1 = vxleft - cleft;
        if (vhead2 <= R(0)) { pm = pright; vxm = vxright; rhom = rhoright; }
        else if (vhead1 >= R(0)) { pm = pleft; vxm = vxleft; rhom = rholeft; }
        else if (vhead2 > R(0) && vtail2 <= R(0)) {
            pm = O::fan(pright, pm, vtail2, vhead2);
            rhom = O::fan(rhoright, rc2, vtail2, vhead2);
            vxm = O::fan(vxright, vxm, vtail2, vhead2);
        } else if (vxm <= R(0) && vtail2 > R(0)) rhom = rc2;
        else if (vxm > R(0) && vtail1 <= R(0)) rhom = rc1;
        else if (vtail1 > R(0) && vhead1 <= R(0)) {
            pm = O::fan(pleft, pm, vtail1, vhead1);
            rhom = O::fan(rholeft, rc1, vtail1, vhead1);
            vxm = O::fan(vxleft, vxm, vtail1, vhead1);
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
            pm = O::fan(pright, pm, vtail2, vhead2);
            rhom = O::fan(rhoright, R(0), vtail2, vhead2);
            vxm = O::fan(vxright, vxm, vtail2, vhead2);
        } else if (vtail1 <= R(0) && vtail2 > R(0)) rhom = R(0);
        else if (vtail1 > R(0) && vhead1 <= R(0)) {
            pm = O::fan(pleft, pm, vtail1, vhead1);
            rhom = O::fan(rholeft, R(0), vtail1, vhead1);
            vxm = O::fan(vxleft, vxm, vtail1, vhead1);
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

// riemannsolver.F90:32-285: classification and Newton iteration in FP64, the
// sampling's costly operations in precision Q (FP64 form: double; mixed
// form: float).
template <class Q = double>
__device__ __forceinline__ Flux riemann_solve(const RScal& Rs, const Face<double>& L, const Face<double>& Rt, int& nit)
{
    // A contact at rest (equal pressures, no normal velocity on either side)
    // is a zero-strength problem: the two-shock branch, its initial guess
    // pcheck = p and one Newton iteration return p* = p and v* = 0 exactly in
    // this arithmetic (p/p, (g1+1)/(g1+1) and 0/(a+b) are exact), and the
    // sampling picks the right state's density.  The transverse sweeps of a
    // planar flow are made of these.  (Not in the reference-order form, whose
    // rhoshock rounds rho (g1+1)/(g1+1) and can differ from rho by an ulp.)
    // The mixed form's FP32 sampling returns the same state from the full
    // path (its shock density is in delta form, exact for pm = p).  Routed
    // through the sampling as pattern 0 so the transverse-velocity upwinding
    // stays one selection of values.
    RSolve S;
    if (L.p == Rt.p && L.u == 0.0 && Rt.u == 0.0) {
        S.rrl = S.rrr = S.rcle = S.cleft = S.rcre = S.cright = 0.0;   // unused by pattern 0
        S.kase = 0;
        S.nit = 1;
        S.pm = L.p;
        S.vxm = 0.0;
    } else {
        riemann_setup(Rs, L, Rt, S);
        if (S.kase == 1 || S.kase == 2) riemann_newton(Rs, L, Rt, S);
    }
    nit = S.nit;
    return riemann_sample<Q>(Rs, L, Rt, S);
}

// ---- conservative update + dual energy (FP64) ------------------------------------------
// Dual-energy selection of integrate.F90:196-216 with the ratio tests
// e_aux/E >= eta cross-multiplied when E > 0 (the reference's division
// otherwise).
__device__ __forceinline__ void dual_energy_d(const IScal& S, double etotnew, double rhonew, double vxnew,
                                              double vynew, double vznew, double eint, double pres,
                                              double& eintnew, double& pnew)
{
    const double vvv = vxnew * vxnew + vynew * vynew + vznew * vznew;
    const double eeaux = etotnew - 0.50 * rhonew * vvv;
    bool ge2, ge1;
    if (etotnew > 0.0) {
        ge2 = eeaux >= S.eta2 * etotnew;
        ge1 = eeaux >= S.eta1 * etotnew;
    } else {
        const double ee1 = eeaux / etotnew;
        ge2 = ee1 >= S.eta2;
        ge1 = ee1 >= S.eta1;
    }
    if (ge2) eintnew = eeaux;
    pnew = (S.gamma - 1) * (ge1 ? eeaux : eintnew);
    if (eintnew <= 0.0) { eintnew = eint; pnew = pres; }
}

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
    dual_energy_d(S, etotnew, rhonew, vxnew, vynew, vznew, eint, pres, eintnew, pnew);
    if (rhonew <= 0.0) rhonew = rho;
    pres = pnew;
    rho = rhonew;
    vx = vxnew;
    vy = vynew;
    vz = vznew;
}

}  // namespace fm
}  // namespace hp
