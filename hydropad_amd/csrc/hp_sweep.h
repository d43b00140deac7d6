// hp_sweep.h -- the sweep kernels, templated on the math policy M (exact
// reference-order arithmetic, or the throughput forms of hp_fast.h) and the
// reconstruction precision R.  Included by hp_kernels_exact.hip (built with
// FP contraction off) and hp_kernels_fast.hip (contraction on).
//
//   sweep_lane_x   x sweep.  One lane per cell of the contiguous axis; the
//                  wave is a 64-cell window over the concatenated x-rows and
//                  neighbours come from DPP wave shifts.  Lanes 3..60 own
//                  their output, so a wave advances 58 cells.
//   sweep_stream   y and z sweeps.  One thread walks a pencil segment along
//                  the strided axis with a register sliding window; lanes of
//                  a wave are consecutive x, so every load/store coalesces.
// Both read the source buffer and write the other one (ping-pong), so pencils
// can be split into segments without a read/write race, and they write the
// periodic face-ghost images of what they store (the one-rank exchange_mesh,
// hydropad/exchange_mesh.F90:47-123) in the same pass.  The last sweep of a
// step also folds in the CFL maximum (hydropad/speed.F90) and, for comoving
// boxes, the expansion step (hydropad/expand.F90).  The per-box bodies x_run
// and stream_run are shared with the persistent batch kernels.
#pragma once
#include <algorithm>
#include <type_traits>
#include "hp_core.h"
#include "hp_fast.h"
#include "hp_kernels.h"

namespace hp {

// Both policies present the same interface to the kernels: cell values,
// interface values and face states are FP64; slopes are of type R (double in
// the FP64 forms, float in the mixed form).
template <class R> struct AllScalars;

// Reference-order arithmetic (hp_core.h); R must be double.
struct ExactMath {
    static constexpr bool G53 = false;
    template <class R> static __device__ __forceinline__ R slope(double am1, double a, double ap1)
    {
        return ajside(am1, a, ap1);
    }
    template <class R> static __device__ __forceinline__ double iface(double am1, double a0, R dm1, R d0)
    {
        return iface_value<double>(am1, a0, dm1, d0);
    }
    static __device__ __forceinline__ bool steep(double a0, double a1, double thr) { return pair_steep(a0, a1, thr); }
    template <class R, int NV, bool GRAV, bool COSMO>
    static __device__ __forceinline__ void cell_faces(const AllScalars<R>& A, const double* a, const double* ivl,
                                                      const double* ivr, double rm2, double rm1, double rp1,
                                                      double rp2, R drm1, R drp1, double pm1, double pp1, bool flat,
                                                      Face<double>& fL, Face<double>& fR, int& nclL, int& nclR);
    template <class R>
    static __device__ __forceinline__ Flux riemann(const RScal& Rs, const Face<double>& L, const Face<double>& Rt,
                                                   int& nit)
    {
        return riemann_solve(Rs, L, Rt, nit);
    }
    static __device__ __forceinline__ void integrate(const IScal& S, const Flux& F1, const Flux& F2, double& p,
                                                     double& rho, double& u, double& v, double& w, double etot,
                                                     double g, bool grav)
    {
        integrate_cell(S, F1, F2, p, rho, u, v, w, etot, g, grav);
    }
    static __device__ __forceinline__ double vsound(double gamma, double p, double rho) { return sqrt(gamma * p / rho); }
};

// Throughput forms (hp_fast.h), delta-form reconstruction.
struct FastMath {
    static constexpr bool G53 = false;
    template <class R> static __device__ __forceinline__ R slope(double am1, double a, double ap1)
    {
        return fm::slope_d<R>(am1, a, ap1);
    }
    template <class R> static __device__ __forceinline__ double iface(double am1, double a0, R dm1, R d0)
    {
        return fm::iface_d<R>(am1, a0, dm1, d0);
    }
    static __device__ __forceinline__ bool steep(double a0, double a1, double thr) { return fm::pair_steep_d(a0, a1, thr); }
    template <class R, int NV, bool GRAV, bool COSMO>
    static __device__ __forceinline__ void cell_faces(const AllScalars<R>& A, const double* a, const double* ivl,
                                                      const double* ivr, double rm2, double rm1, double rp1,
                                                      double rp2, R drm1, R drp1, double pm1, double pp1, bool flat,
                                                      Face<double>& fL, Face<double>& fR, int& nclL, int& nclR);
    // the mixed form (R = float) samples the Riemann solution at single-precision accuracy
    template <class R>
    static __device__ __forceinline__ Flux riemann(const RScal& Rs, const Face<double>& L, const Face<double>& Rt,
                                                   int& nit)
    {
        return fm::riemann_solve<R>(Rs, L, Rt, nit);
    }
    static __device__ __forceinline__ void integrate(const IScal& S, const Flux& F1, const Flux& F2, double& p,
                                                     double& rho, double& u, double& v, double& w, double etot,
                                                     double g, bool grav)
    {
        fm::integrate_cell(S, F1, F2, p, rho, u, v, w, etot, g, grav);
    }
    // the CFL sound speed keeps speed.F90's IEEE sqrt(gamma*p/rho), so the
    // fused maximum equals hp_speed's bit for bit and every path gets the same dt
    static __device__ __forceinline__ double vsound(double gamma, double p, double rho) { return sqrt(gamma * p / rho); }
};

// The throughput forms with gamma = 5/3 fixed at compile time: the
// gamma-derived scalars become constants (same values, bit for bit, as
// hp_params_init computes them from 5.0/3.0), the Riemann solver's powers are
// the root/product forms only, and the freed registers raise occupancy.
struct FastMath53 : FastMath {
    static constexpr bool G53 = true;
};

// field index (0 p, 1 rho, 2 vx, 3 vy, 4 vz) of the normal velocity and the
// transverse pair, hydro/ppmsolver.F90:105-107 (x), :227-229 (y), :355-357 (z)
template <int DIR> struct Dirmap;
template <> struct Dirmap<1> { static constexpr int U = 2, V = 3, W = 4; };
template <> struct Dirmap<2> { static constexpr int U = 3, V = 2, W = 4; };
template <> struct Dirmap<3> { static constexpr int U = 4, V = 3, W = 2; };

// total(p,rho,vx,vy,vz) with the physical components in x,y,z order
// (hydro/total.F90:15; ppmsolver.F90:149,277,405 pass vx3d,vy3d,vz3d).
template <int DIR>
__device__ __forceinline__ double total_phys(double rgamma1, double p, double r, double u, double v, double w)
{
    double vx, vy, vz;
    if (DIR == 1) { vx = u; vy = v; vz = w; }
    else if (DIR == 2) { vx = v; vy = u; vz = w; }
    else { vx = w; vy = v; vz = u; }
    return p * rgamma1 + 0.5 * r * (vx * vx + vy * vy + vz * vz);
}

// ---- DPP whole-wave shifts (gfx9 wave_shr:1 / wave_shl:1) -------------------
constexpr int DPP_WAVE_SHL1 = 0x130;   // lane l receives lane l+1
constexpr int DPP_WAVE_SHR1 = 0x138;   // lane l receives lane l-1

template <int CTRL> __device__ __forceinline__ int dpp_i(int x)
{
    return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ double dpp(double x)
{
    int lo = dpp_i<CTRL>(__double2loint(x));
    int hi = dpp_i<CTRL>(__double2hiint(x));
    return __hiloint2double(hi, lo);
}
template <int CTRL> __device__ __forceinline__ float dpp(float x)
{
    return __int_as_float(dpp_i<CTRL>(__float_as_int(x)));
}
template <int CTRL> __device__ __forceinline__ int dpp(int x) { return dpp_i<CTRL>(x); }
template <class T> __device__ __forceinline__ T from_left(T x) { return dpp<DPP_WAVE_SHR1>(x); }
template <class T> __device__ __forceinline__ T from_right(T x) { return dpp<DPP_WAVE_SHL1>(x); }

template <class R> __device__ __forceinline__ Face<R> from_left(const Face<R>& f)
{
    Face<R> o;
    o.p = from_left(f.p); o.r = from_left(f.r); o.u = from_left(f.u);
    o.v = from_left(f.v); o.w = from_left(f.w);
    return o;
}
__device__ __forceinline__ Flux from_right(const Flux& f)
{
    Flux o;
    o.p = from_right(f.p); o.r = from_right(f.r); o.u = from_right(f.u);
    o.v = from_right(f.v); o.w = from_right(f.w);
    return o;
}

// ---- per-box scalars ----------------------------------------------------------
template <class R> struct AllScalars {
    SweepScalars<R> S;
    RScal Rs;
    IScal Is;
    double rgamma1, gamma, mingradflat;
};

// gamma-derived constants for gamma = 5/3, by the expressions of
// hydropad/initialize_parameters.F90:64-69 (and hp_params_init)
struct Gamma53 {
    static constexpr double gamma = 5.0 / 3.0;
    static constexpr double gamma1 = (gamma + 1.0) / (gamma - 1.0);
    static constexpr double rgamma = 1.0 / gamma;
    static constexpr double rgamma1 = 1.0 / (gamma - 1);
    static constexpr double m = (gamma - 1.0) / (2.0 * gamma);
    static constexpr double rm = 1.0 / m;
    static constexpr double gf = (gamma + 1.0) / (2.0 * gamma);
};

template <class M, class R>
__device__ __forceinline__ AllScalars<R> load_scalars(const Params* __restrict__ prm, const DevStatus* st)
{
    const Params P = *prm;
    const int cap = st->newton_cap;
    double gamma = P.gamma, gamma1 = P.gamma1, rgamma = P.rgamma, rgamma1 = P.rgamma1;
    double m = P.m, rm = P.rm, gf = P.gf;
    int m5 = fabs(P.m - 0.2) < 1e-15;
    if constexpr (M::G53) {
        gamma = Gamma53::gamma; gamma1 = Gamma53::gamma1; rgamma = Gamma53::rgamma; rgamma1 = Gamma53::rgamma1;
        m = Gamma53::m; rm = Gamma53::rm; gf = Gamma53::gf;
        m5 = 1;
    }
    AllScalars<R> A;
    A.S.gamma = (R)gamma; A.S.dt = (R)P.dt; A.S.rat = (R)P.rat; A.S.dx = (R)P.dx;
    A.S.rdx = (R)P.rdx; A.S.dat = (R)P.dat; A.S.at = (R)P.at;
    A.S.flatvalue = (R)P.flatvalue; A.S.mingradflat = (R)P.mingradflat;
    A.S.dtrat = (R)(P.dt * P.rat);
    A.S.cdat2 = (R)(2.0 * P.dat * P.rat * P.dt);
    A.Rs.gamma = gamma; A.Rs.gamma1 = gamma1; A.Rs.rgamma = rgamma; A.Rs.rgamma1 = rgamma1;
    A.Rs.m = m; A.Rs.rm = rm; A.Rs.gf = gf;
    A.Rs.dadt = 1.0 - 0.5 * P.dat * P.dt * P.rat;     // riemannsolver.F90:36
    A.Rs.dmax = P.dmax;
    A.Rs.m5 = m5;
    A.Rs.nmax = cap > 0 ? cap : K_NNEWTMAX;
    A.Rs.nskip = K_NNEWTMAX - A.Rs.nmax;
    A.Is.rdtath = P.rdtath; A.Is.rdx = P.rdx; A.Is.gamma = gamma; A.Is.rgamma1 = rgamma1;
    A.Is.eta1 = P.eta1; A.Is.eta2 = P.eta2;
    A.rgamma1 = rgamma1; A.gamma = gamma; A.mingradflat = P.mingradflat;
    return A;
}

// Parabolas of one cell and its two traced face states, reference order.
//   a[]            the cell's values (p, rho, u, v, w[, g])
//   ivl[], ivr[]   interface values before limiting at its left / right interface
//   rm2..rp2       rho of cells c-2..c+2 (steepening), drm1/drp1 rho slopes of c-1, c+1
//   pm1, pp1       p of cells c-1, c+1
template <class R, int NV, bool GRAV, bool COSMO>
__device__ __forceinline__ void ExactMath::cell_faces(const AllScalars<R>& A, const double* a, const double* ivl,
                                                      const double* ivr, double rm2, double rm1, double rp1,
                                                      double rp2, R drm1, R drp1, double pm1, double pp1, bool flat,
                                                      Face<double>& fL, Face<double>& fR, int& nclL, int& nclR)
{
    const SweepScalars<double>& S = A.S;
    const double flatv = flat ? S.flatvalue : 0.0;
    const double addp = fabs(pp1 - pm1) / tmin(pm1, pp1);                          // order2interp.F90:152
    const double addrho = 0.50 * S.gamma * fabs(rp1 - rm1) / tmin(rm1, rp1);       // :153
    Parab<double> q[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double ajl = ivl[v], ajr = ivr[v];
        if (v == VR && addrho >= addp) steepen(rm2, rm1, a[VR], rp1, rp2, drm1, drp1, S.dx, ajl, ajr);
        q[v] = limit_parab(a[v], ajl, ajr, flatv);
    }
    const double cs = sqrt(S.gamma * a[VP] / a[VR]);                               // ppmsolver.F90:109
    fR = trace_right_face<double, GRAV, COSMO>(S, q, a[VU], cs, nclR);
    fL = trace_left_face<double, GRAV, COSMO>(S, q, a[VU], cs, nclL);
}

template <class R, int NV, bool GRAV, bool COSMO>
__device__ __forceinline__ void FastMath::cell_faces(const AllScalars<R>& A, const double* a, const double* ivl,
                                                     const double* ivr, double rm2, double rm1, double rp1,
                                                     double rp2, R drm1, R drp1, double pm1, double pp1, bool flat,
                                                     Face<double>& fL, Face<double>& fR, int& nclL, int& nclR)
{
    fm::cell_faces_d<R, NV, GRAV, COSMO>(A.S, A.gamma, a, ivl, ivr, rm2, rm1, rp1, rp2, drm1, drp1, pm1, pp1, flat,
                                         fL, fR, nclL, nclR);
}

// Store one updated cell and, for the axes in `mask`, its periodic face-ghost
// image (ghost g <- cell g+N on the low side, ghost N+nb+g <- cell nb+g).
// Each image is its own conditional store: only cells within 2 nb of a face
// have one, so a wave almost always skips them (a list of shifts built per
// cell instead cost every cell some sixty instructions of bookkeeping).
__device__ __forceinline__ void store_cell(double* const* out, long q, const double* val, int ix, int iy, int iz,
                                           int mask, int nx, int ny, int nz, int nb)
{
#pragma unroll
    for (int f = 0; f < 5; ++f) out[f][q] = val[f];
    auto image = [&](long shift) {
#pragma unroll
        for (int f = 0; f < 5; ++f) out[f][q + shift] = val[f];
    };
    if (mask & 1) {
        const int N = nx - 2 * nb;
        if (ix < 2 * nb) image(N);
        if (ix >= N) image(-N);
    }
    if (mask & 2) {
        const long N = ny - 2 * nb;
        if (iy < 2 * nb) image(N * nx);
        if (iy >= N) image(-N * nx);
    }
    if (mask & 4) {
        const long N = nz - 2 * nb;
        if (iz < 2 * nb) image(N * nx * (long)ny);
        if (iz >= N) image(-N * nx * (long)ny);
    }
}

// What a sweep does after integrate, besides the store: the expansion step of
// hydropad/expand.F90:18-41 against the start-of-step state (cosmological
// boxes, last sweep of a step; old[0] == nullptr when off) and the CFL maximum.
struct Epilogue {
    const double* old[4];    // start-of-step p, vx, vy, vz
    const double* fac;       // device (afact, afact1, afact2), read by load_factors
    double afact, afact1, afact2;
    int ghost_mask;          // bit d-1: write the periodic face-ghost images along axis d
    int do_speed;
};

// Per-lane tallies, reduced once per wave.
struct Tally {
    int nclamp, nitmax, nonconv;
    double vmax;
};

// Read the expansion factors of the step from device memory (kernel entry).
__device__ __forceinline__ void load_factors(Epilogue& E)
{
    if (E.old[0]) { E.afact = E.fac[0]; E.afact1 = E.fac[1]; E.afact2 = E.fac[2]; }
}

__device__ __forceinline__ Tally tally0()
{
    Tally t;
    t.nclamp = 0; t.nitmax = 0; t.nonconv = 0; t.vmax = 0.0;
    return t;
}

// expand.F90:22-38 on one cell; val[] is (p, rho, vx, vy, vz) in physical order.
// afact1 = 1/(1+afact), afact2 = 1/(2+afact) come precomputed (expand.F90:19-20).
__device__ __forceinline__ void expand_cell(const Epilogue& E, long q, double* val)
{
    const double po = E.old[0][q];
#pragma unroll
    for (int f = 2; f < 5; ++f) val[f] = (2.0 * val[f] - E.afact * E.old[f - 1][q]) * E.afact2;
    const double ppp = val[0];
    double pn = (val[0] - E.afact * po) * E.afact1;
    if (pn <= 0.0) pn = (ppp > 0.0) ? ppp : po;
    val[0] = pn;
}

// The epilogue of one updated cell before its store: expansion, CFL maximum.
template <class M, bool EXP>
__device__ __forceinline__ void settle_cell(const Epilogue& E, double gamma, long q, double* val, Tally& T)
{
    if (EXP) expand_cell(E, q, val);
    if (E.do_speed) {
        const double vs = M::vsound(gamma, val[0], val[1]);
        T.vmax = tmax(T.vmax, tmax(tmax(fabs(val[2]), fabs(val[3])), fabs(val[4])) + vs);
    }
}

// The epilogue of one updated cell: expansion, store (+ ghost images), CFL maximum.
template <class M, bool EXP>
__device__ __forceinline__ void finish_cell(const BoxView& B, const Epilogue& E, double gamma, long q, double* val,
                                            int ix, int iy, int iz, Tally& T)
{
    settle_cell<M, EXP>(E, gamma, q, val, T);
    store_cell(B.out, q, val, ix, iy, iz, E.ghost_mask, B.nx, B.ny, B.nz, B.nb);
}

// Wave-level reduction of the per-lane counters, then one atomic per wave.
__device__ __forceinline__ void flush_counters(DevStatus* st, const Tally& T, bool want_speed)
{
    int nitmax = T.nitmax, nonconv = T.nonconv, nclamp = T.nclamp;
    double vmax = T.vmax;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        nitmax = max(nitmax, __shfl_xor(nitmax, off, 64));
        nonconv += __shfl_xor(nonconv, off, 64);
        nclamp += __shfl_xor(nclamp, off, 64);
        if (want_speed) vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (nitmax > 0) atomicMax(&st->newton_max, nitmax);
        if (nonconv) atomicAdd(&st->newton_nonconv, (unsigned long long)nonconv);
        if (nclamp) atomicAdd(&st->clamps, (unsigned long long)nclamp);
        if (want_speed) atomicMax(&st->velmax_bits, (unsigned long long)__double_as_longlong(vmax));
    }
}

// Block-level form for the single-box kernels: the waves' partial results
// meet in LDS and one thread issues the atomics, so a launch of many short
// waves does not serialise on the status block.  Every wave of the block
// must call it (no early return before it).
__device__ __forceinline__ void flush_counters_block(DevStatus* st, const Tally& T, bool want_speed)
{
    __shared__ int s_nit[4], s_nc[4], s_cl[4];
    __shared__ double s_vm[4];
    int nitmax = T.nitmax, nonconv = T.nonconv, nclamp = T.nclamp;
    double vmax = T.vmax;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        nitmax = max(nitmax, __shfl_xor(nitmax, off, 64));
        nonconv += __shfl_xor(nonconv, off, 64);
        nclamp += __shfl_xor(nclamp, off, 64);
        if (want_speed) vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_nit[w] = nitmax;
        s_nc[w] = nonconv;
        s_cl[w] = nclamp;
        s_vm[w] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int k = 1; k < nw; ++k) {
            nitmax = max(nitmax, s_nit[k]);
            nonconv += s_nc[k];
            nclamp += s_cl[k];
            vmax = fmax(vmax, s_vm[k]);
        }
        if (nitmax > 0) atomicMax(&st->newton_max, nitmax);
        if (nonconv) atomicAdd(&st->newton_nonconv, (unsigned long long)nonconv);
        if (nclamp) atomicAdd(&st->clamps, (unsigned long long)nclamp);
        if (want_speed) atomicMax(&st->velmax_bits, (unsigned long long)__double_as_longlong(vmax));
    }
}

// Newton bookkeeping of one interface.
__device__ __forceinline__ void count_newton(int nit, int& nitmax, int& nonconv)
{
    if (nit < 0) ++nonconv;
    else nitmax = max(nitmax, nit);
}

// ============================================================================
// x sweep: lane per cell, DPP neighbours.  A wave advances chunks [c0, c1) of
// the concatenated x-rows of box B, 58 output cells per chunk.
__device__ __forceinline__ unsigned x_chunks(int nx, int ny, int nz, int nb)
{
    const unsigned total = (unsigned)(ny - 2 * nb) * (unsigned)(nz - 2 * nb) * (unsigned)nx;
    return (total + 57u) / 58u;
}

template <class M, class R, bool GRAV, bool COSMO, bool EXP>
__device__ __forceinline__ void x_run(const BoxView& B, const AllScalars<R>& A, const Epilogue& E, unsigned c0,
                                      unsigned c1, Tally& T)
{
    using D = Dirmap<1>;
    constexpr int NV = GRAV ? 6 : 5;
    const int nx = B.nx, ny = B.ny, nb = B.nb;
    const int Nx = nx - 2 * nb;
    const unsigned Ny = (unsigned)(ny - 2 * nb);
    const unsigned total = Ny * (unsigned)(B.nz - 2 * nb) * (unsigned)nx;   // concatenated x-rows
    const int lane = threadIdx.x & 63;

    // cell of this lane in chunk c: virtual position 58c + lane - 3
    auto locate = [&](unsigned c, long& q, int& i, int& j, int& k, bool& in_range) {
        const int v = (int)(58u * c) + lane - 3;
        in_range = v >= 0 && (unsigned)v < total;
        const unsigned vc = v < 0 ? 0u : ((unsigned)v >= total ? total - 1u : (unsigned)v);
        const unsigned row = vc / (unsigned)nx;
        i = (int)(vc - row * (unsigned)nx);
        const unsigned kz = row / Ny;
        j = (int)(row - kz * Ny) + nb;
        k = (int)kz + nb;
        q = i + (long)nx * (j + (long)ny * k);
    };
    auto load = [&](long q, double* cd) {
        cd[VP] = B.in[0][q]; cd[VR] = B.in[1][q];
        cd[VU] = B.in[D::U][q]; cd[VV] = B.in[D::V][q]; cd[VW] = B.in[D::W][q];
        if (GRAV) cd[NV - 1] = -B.gforce[3 * q + 0];
    };

    // The FP64 form prefetches the next chunk into registers; the mixed form
    // does where the ten registers still leave four waves per SIMD (the
    // gamma = 5/3 kernel without gravity or expansion terms, 125 VGPRs);
    // elsewhere its ten fewer registers give it a wave more per SIMD, which
    // hides the loads better than the prefetch does.
    constexpr bool PREFETCH = std::is_same<R, double>::value || (M::G53 && !GRAV && !COSMO);
    long qn = 0;
    int in_ = 0, jn = 0, kn = 0;
    bool rn = false;
    double nxt[NV];
    if constexpr (PREFETCH) {
        locate(c0, qn, in_, jn, kn, rn);
        load(qn, nxt);
    }
    for (unsigned c = c0; c < c1; ++c) {
        long q;
        int i, j, k;
        bool in_range;
        double cd[NV];
        if constexpr (PREFETCH) {
            q = qn; i = in_; j = jn; k = kn; in_range = rn;
#pragma unroll
            for (int s = 0; s < NV; ++s) cd[s] = nxt[s];
            if (c + 1 < c1) {                          // prefetch the next chunk
                locate(c + 1, qn, in_, jn, kn, rn);
                load(qn, nxt);
            }
        } else {
            locate(c, q, i, j, k, in_range);
            load(q, cd);
        }

        double am1[NV], ap1[NV], iv[NV], ivr[NV];
        R d[NV], dm1[NV];
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            am1[s] = from_left(cd[s]);
            ap1[s] = from_right(cd[s]);
        }
#pragma unroll
        for (int s = 0; s < NV; ++s) d[s] = M::template slope<R>(am1[s], cd[s], ap1[s]);   // slope of this cell
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            dm1[s] = from_left(d[s]);
            iv[s] = M::iface(am1[s], cd[s], dm1[s], d[s]);                // interface c-1/2
        }
#pragma unroll
        for (int s = 0; s < NV; ++s) ivr[s] = from_right(iv[s]);          // interface c+1/2

        // flattening: pair bit of (c-1,c), OR over pairs c-2..c+2 via the wave mask
        bool bpair = false;
#pragma unroll
        for (int s = 0; s < 5; ++s) bpair |= M::steep(am1[s], cd[s], A.mingradflat);
        const uint64_t Mb = __ballot(bpair);
        const uint64_t Mf = (Mb << 1) | Mb | (Mb >> 1) | (Mb >> 2);
        const bool flat = (A.mingradflat <= 0.0) || ((Mf >> lane) & 1ull);

        const double rm2 = from_left(am1[VR]);
        const double rp2 = from_right(ap1[VR]);
        const R drp1 = from_right(d[VR]);
        int nclL = 0, nclR = 0;
        Face<double> fL, fR;
        M::template cell_faces<R, NV, GRAV, COSMO>(A, cd, iv, ivr, rm2, am1[VR], ap1[VR], rp2, dm1[VR], drp1,
                                                   am1[VP], ap1[VP], flat, fL, fR, nclL, nclR);

        // Riemann problem at interface c-1/2: right face of c-1 | left face of c;
        // the updates of lanes 3..60 use the interfaces of lanes 3..61
        const Face<double> fRm1 = from_left(fR);
        const bool solve = in_range && lane >= 3 && lane <= 61;
        int nit = 0;
        Flux F;
        if (solve) {
            F = M::template riemann<R>(A.Rs, fRm1, fL, nit);
        } else {
            F.p = F.r = F.u = F.v = F.w = 0.0;
        }
        // high-density-peak floor (riemannsolver.F90:279-285), needs rho c-3..c
        const uint64_t M1 = __ballot(cd[VR] > A.Rs.dmax * am1[VR]);
        const uint64_t M2 = __ballot(am1[VR] > A.Rs.dmax * cd[VR]);
        const uint64_t Mc = (M1 << 2) & M2;
        if (Mc) {
            const double r1 = from_left(rm2);
            const double u3 = am1[VU], u2 = from_left(am1[VU]);
            if ((Mc >> lane) & 1ull) peak_floor(A.Rs, F, r1, rm2, am1[VR], cd[VR], u2, u3);
        }
        const Flux F2 = from_right(F);                                    // interface c+1/2
        // counters per interface, as the reference's loop over interfaces
        // nb-1/2 .. nb+N-1/2 (ppmsolver.F90:98-138): interface c-1/2 takes the
        // clamps of c-1's right face and c's left face; the last interior
        // cell also counts interface c+1/2, solved by the lane to its right
        const int nclRm1 = from_left(nclR), nclLp1 = from_right(nclL), nitp1 = from_right(nit);

        const bool valid = in_range && lane >= 3 && lane <= 60 && i >= nb && i < nb + Nx;
        // Take delivery of the next chunk's prefetch (issued a whole chunk ago)
        // here, before this chunk's stores: vmcnt counts stores as well as
        // loads, so a first use at the top of the next chunk would also wait
        // for these stores to complete.  The empty asm is that use.
        if constexpr (PREFETCH) {
#pragma unroll
            for (int s = 0; s < NV; ++s) asm volatile("" : "+v"(nxt[s]));
        }
        if (valid) {
            count_newton(nit, T.nitmax, T.nonconv);
            T.nclamp += nclRm1 + nclL;
            if (i == nb + Nx - 1) {
                count_newton(nitp1, T.nitmax, T.nonconv);
                T.nclamp += nclR + nclLp1;
            }
            double pres = cd[VP], rho = cd[VR], u = cd[VU], vv = cd[VV], w = cd[VW];
            const double etot = total_phys<1>(A.rgamma1, pres, rho, u, vv, w);
            M::integrate(A.Is, F, F2, pres, rho, u, vv, w, etot, GRAV ? cd[NV - 1] : 0.0, GRAV);
            double val[5];
            val[0] = pres; val[1] = rho; val[D::U] = u; val[D::V] = vv; val[D::W] = w;
            finish_cell<M, EXP>(B, E, A.gamma, q, val, i, j, k, T);
        }
    }
}

template <class M, class R, bool GRAV, bool COSMO, bool EXP>
__global__ void __launch_bounds__(256) sweep_lane_x(BoxView B, Epilogue E, int chunks_per_wave)
{
    if (EXP) load_factors(E);
    const unsigned nchunks = x_chunks(B.nx, B.ny, B.nz, B.nb);
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned c0 = wave * (unsigned)chunks_per_wave;
    const unsigned c1 = min(c0 + (unsigned)chunks_per_wave, nchunks);
    Tally T = tally0();
    if (c0 < c1) {                                     // waves past the end only join the flush
        const AllScalars<R> A = load_scalars<M, R>(B.prm, B.st);
        x_run<M, R, GRAV, COSMO, EXP>(B, A, E, c0, c1, T);
    }
    flush_counters_block(B.st, T, E.do_speed != 0);
}

// ============================================================================
// y / z sweeps: one thread streams along segment `seg` (cells
// [seg*seglen, min((seg+1)*seglen, N))) of pencil `pen` of box B.
template <int DIR> __device__ __forceinline__ void stream_geometry(const BoxView& B, int& Nax, int& Nbb, long& s_ax,
                                                                  long& s_b)
{
    const int nb = B.nb;
    if (DIR == 2) { Nax = B.ny - 2 * nb; Nbb = B.nz - 2 * nb; s_ax = B.nx; s_b = (long)B.nx * B.ny; }
    else { Nax = B.nz - 2 * nb; Nbb = B.ny - 2 * nb; s_ax = (long)B.nx * B.ny; s_b = B.nx; }
}

template <class M, class R, int DIR, bool GRAV, bool COSMO, bool EXP>
__device__ __forceinline__ void stream_run(const BoxView& B, const AllScalars<R>& A, const Epilogue& E, long pen,
                                           int seg, int seglen, bool active, Tally& T)
{
    using D = Dirmap<DIR>;
    constexpr int NV = GRAV ? 6 : 5;
    const int nx = B.nx, nb = B.nb;
    const int Nx = nx - 2 * nb;
    int Nax, Nbb;
    long s_ax, s_b;
    stream_geometry<DIR>(B, Nax, Nbb, s_ax, s_b);
    const unsigned upen = (unsigned)pen;                 // pencils per box < 2^31
    const int ib = (int)(upen / (unsigned)Nx), ia = (int)(upen - (unsigned)ib * (unsigned)Nx);
    const int s0 = seg * seglen;
    const int s1 = min(s0 + seglen, Nax);
    const long base = (nb + ia) + (long)(nb + ib) * s_b + (long)nb * s_ax;   // interior cell 0 of the pencil
    const bool flat_always = A.mingradflat <= 0.0;

    double w0[NV], w1[NV], w2[NV], w3[NV];   // cells k-4 .. k-1
    double r5 = 0.0;                         // rho of cell k-5
    R d1[NV], d2[NV];                        // slopes of cells k-3, k-2
    double iv2[NV];                          // interface value at k-2 (left of cell k-2)
    unsigned bits = 0;                       // pair bits of k-3, k-2, k-1
    Face<double> fRprev;                     // right face of cell k-3
    Flux Fprev;                              // flux at interface k-3 (left of cell k-3)
#pragma unroll
    for (int s = 0; s < NV; ++s) { w0[s] = w1[s] = w2[s] = w3[s] = iv2[s] = 0.0; d1[s] = d2[s] = R(0); }
    fRprev.p = fRprev.r = fRprev.u = fRprev.v = fRprev.w = 0.0;
    Fprev.p = Fprev.r = Fprev.u = Fprev.v = Fprev.w = 0.0;

    // The stores of the cell updated in one iteration are issued at the top of
    // the next, before that iteration's prefetch: the wait for the prefetched
    // values (vmcnt counts stores as well as loads) then finds both a whole
    // iteration old instead of stalling on stores issued just before it.  The
    // comoving kernels keep the store in place (the five values held across
    // the iteration boundary would cost them occupancy or spills) and take
    // delivery of the prefetch before it instead.
    constexpr bool DEFER = !COSMO;
    double pend[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int ixc = nb + ia, ibc = nb + ib;
    auto store_pending = [&](int kc) {                 // cell kc of the segment, updated last iteration
        const int ic = nb + kc;
        if (DIR == 2) store_cell(B.out, base + (long)kc * s_ax, pend, ixc, ic, ibc, E.ghost_mask, B.nx, B.ny, B.nz, nb);
        else store_cell(B.out, base + (long)kc * s_ax, pend, ixc, ibc, ic, E.ghost_mask, B.nx, B.ny, B.nz, nb);
    };

    if (active) {
        double nxt[NV];
        {
            const long q0 = base + (long)(s0 - 3) * s_ax;
            nxt[VP] = B.in[0][q0]; nxt[VR] = B.in[1][q0];
            nxt[VU] = B.in[D::U][q0]; nxt[VV] = B.in[D::V][q0]; nxt[VW] = B.in[D::W][q0];
            if (GRAV) nxt[NV - 1] = -B.gforce[3 * q0 + (DIR - 1)];
        }
        if constexpr (!DEFER) {     // the first cell too, so the loop head holds no wait
#pragma unroll
            for (int s = 0; s < NV; ++s) asm volatile("" : "+v"(nxt[s]));
        }
        for (int kk = s0 - 3; kk <= s1 + 2; ++kk) {
            const long q = base + (long)kk * s_ax;
            double cur[NV];
#pragma unroll
            for (int s = 0; s < NV; ++s) cur[s] = nxt[s];
            if (DEFER && kk >= s0 + 4) store_pending(kk - 4);
            if (kk + 1 <= s1 + 2) {                    // prefetch the next cell of the pencil
                const long qn = q + s_ax;
                nxt[VP] = B.in[0][qn]; nxt[VR] = B.in[1][qn];
                nxt[VU] = B.in[D::U][qn]; nxt[VV] = B.in[D::V][qn]; nxt[VW] = B.in[D::W][qn];
                if (GRAV) nxt[NV - 1] = -B.gforce[3 * qn + (DIR - 1)];
            }

            R d3[NV];
            double iv1[NV];
            bool bk = false;
            if (kk >= s0 - 1) {
#pragma unroll
                for (int s = 0; s < NV; ++s) d3[s] = M::template slope<R>(w2[s], w3[s], cur[s]);
            } else {
#pragma unroll
                for (int s = 0; s < NV; ++s) d3[s] = R(0);
            }
            if (kk >= s0 - 2) {
#pragma unroll
                for (int s = 0; s < 5; ++s) bk |= M::steep(w3[s], cur[s], A.mingradflat);
            }
#pragma unroll
            for (int s = 0; s < NV; ++s) iv1[s] = M::iface(w2[s], w3[s], d2[s], d3[s]);

            Face<double> fL = fRprev, fR = fRprev;
            const int ic = kk - 2;                                   // cell k-2, interface at its left
            if (kk >= s0 + 1) {     // parabola and faces of cell k-2
                // clamp count: both faces of the cells s0-1 .. s1 a segment
                // reconstructs (per interface only in the x sweep: an exact
                // count here costs the FP64 form half its occupancy)
                const bool flat = flat_always || bits != 0u || bk;
                M::template cell_faces<R, NV, GRAV, COSMO>(A, w2, iv2, iv1, w0[VR], w1[VR], w3[VR], cur[VR], d1[VR],
                                                           d3[VR], w1[VP], w3[VP], flat, fL, fR, T.nclamp, T.nclamp);
            }
            Flux F = Fprev;
            if (kk >= s0 + 2) {     // Riemann problem at interface k-2
                int nit;
                F = M::template riemann<R>(A.Rs, fRprev, fL, nit);
                peak_floor(A.Rs, F, r5, w0[VR], w1[VR], w2[VR], w0[VU], w1[VU]);
                // Newton count once per interface of the pencil (0 .. Nax), as
                // the reference's loop over interfaces: a segment owns
                // interfaces s0 .. s1-1 and the pencil's last segment also Nax
                if (kk != s1 + 2 || s1 == Nax) count_newton(nit, T.nitmax, T.nonconv);
            }
            if constexpr (!DEFER) {
                // take delivery of the prefetch before this cell's stores
                // (the empty asm is the use; see x_run)
#pragma unroll
                for (int s = 0; s < NV; ++s) asm volatile("" : "+v"(nxt[s]));
            }
            if (kk >= s0 + 3) {     // update cell k-3
                double pres = w1[VP], rho = w1[VR], u = w1[VU], vv = w1[VV], w = w1[VW];
                const double etot = total_phys<DIR>(A.rgamma1, pres, rho, u, vv, w);
                M::integrate(A.Is, Fprev, F, pres, rho, u, vv, w, etot, GRAV ? w1[NV - 1] : 0.0, GRAV);
                pend[0] = pres; pend[1] = rho; pend[D::U] = u; pend[D::V] = vv; pend[D::W] = w;
                settle_cell<M, EXP>(E, A.gamma, q - 3 * s_ax, pend, T);
                if (!DEFER) store_pending(kk - 3);
            }
            // slide the window
            r5 = w0[VR];
#pragma unroll
            for (int s = 0; s < NV; ++s) {
                w0[s] = w1[s]; w1[s] = w2[s]; w2[s] = w3[s]; w3[s] = cur[s];
                d1[s] = d2[s]; d2[s] = d3[s]; iv2[s] = iv1[s];
            }
            bits = (bits >> 1) | ((bk ? 1u : 0u) << 2);
            fRprev = fR;
            Fprev = F;
        }
        if (DEFER && s1 + 2 >= s0 + 3) store_pending(s1 - 1);   // the last cell, updated at kk = s1 + 2
    }
}

template <class M, class R, int DIR, bool GRAV, bool COSMO, bool EXP>
__global__ void __launch_bounds__(256) sweep_stream(BoxView B, Epilogue E, int seglen, int nseg)
{
    if (EXP) load_factors(E);
    int Nax, Nbb;
    long s_ax, s_b;
    stream_geometry<DIR>(B, Nax, Nbb, s_ax, s_b);
    const long npen = (long)(B.nx - 2 * B.nb) * Nbb;
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;   // items per sweep < 2^31
    const bool active = (long)tid < npen * nseg;
    const unsigned tt = active ? tid : 0u;
    const int seg = (int)(tt / (unsigned)npen);
    const long pen = (long)(tt - (unsigned)seg * (unsigned)npen);
    const AllScalars<R> A = load_scalars<M, R>(B.prm, B.st);
    Tally T = tally0();
    stream_run<M, R, DIR, GRAV, COSMO, EXP>(B, A, E, pen, seg, seglen, active, T);
    flush_counters_block(B.st, T, E.do_speed != 0);
}

// host launcher of one sweep
template <class M, class R, bool G, bool C, bool X>
static hipError_t launch_sweep_t(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    Epilogue E;
    for (int f = 0; f < 4; ++f) E.old[f] = L.old[f];
    E.fac = L.fac;
    E.afact = E.afact1 = E.afact2 = 0.0;
    E.ghost_mask = L.ghost_mask;
    E.do_speed = L.do_speed;
    const int nb = b.nb;
    const int Nx = b.nx - 2 * nb, Ny = b.ny - 2 * nb, Nz = b.nz - 2 * nb;
    const int block = L.block > 0 ? L.block : 256;
    if (L.dir == 1) {
        const long total = (long)Ny * Nz * b.nx;
        const long nchunks = (total + 57) / 58;
        // about 16 resident rounds of waves over 256 CUs x 4 SIMDs; each wave
        // walks its run of chunks with a one-chunk register prefetch
        // waves of 16 chunks (fewer for small boxes, to keep ~16K waves):
        // short runs balance the divergent Riemann work across the chip
        long cpw = std::min<long>(16, nchunks / 16384);
        if (L.xcpw > 0) cpw = L.xcpw;
        if (cpw < 1) cpw = 1;
        const long waves = (nchunks + cpw - 1) / cpw;
        const long threads = waves * 64;
        const long grid = (threads + block - 1) / block;
        hipLaunchKernelGGL((sweep_lane_x<M, R, G, C, X>), dim3((unsigned)grid), dim3(block), 0, s, b, E, (int)cpw);
    } else {
        const int Nax = (L.dir == 2) ? Ny : Nz;
        const int Nbb = (L.dir == 2) ? Nz : Ny;
        // default: split pencils into segments of at least 64 cells until there
        // are ~16K waves (each segment recomputes a 6-cell halo)
        int seglen = L.seglen;
        if (seglen <= 0) {
            const long npen = (long)Nx * ((L.dir == 2) ? Nz : Ny);
            seglen = (int)std::min<long>(Nax, std::max<long>(64, (long)Nax * npen / (16384L * 64)));
        }
        if (seglen > Nax) seglen = Nax;
        const int nseg = (Nax + seglen - 1) / seglen;
        const long threads = (long)Nx * Nbb * nseg;
        const long grid = (threads + block - 1) / block;
        if (L.dir == 2)
            hipLaunchKernelGGL((sweep_stream<M, R, 2, G, C, X>), dim3((unsigned)grid), dim3(block), 0, s, b, E,
                               seglen, nseg);
        else
            hipLaunchKernelGGL((sweep_stream<M, R, 3, G, C, X>), dim3((unsigned)grid), dim3(block), 0, s, b, E,
                               seglen, nseg);
    }
    return hipGetLastError();
}

// dispatch over gravity / comoving / fused expansion (expansion only with comoving)
template <class M, class R>
static hipError_t launch_sweep_variants(const BoxView& b, const SweepLaunch& L, hipStream_t s)
{
    const bool x = L.old[0] != nullptr;
    if (x && !L.cosmo) return hipErrorInvalidValue;
    if (L.grav) {
        if (!L.cosmo) return launch_sweep_t<M, R, true, false, false>(b, L, s);
        return x ? launch_sweep_t<M, R, true, true, true>(b, L, s) : launch_sweep_t<M, R, true, true, false>(b, L, s);
    }
    if (!L.cosmo) return launch_sweep_t<M, R, false, false, false>(b, L, s);
    return x ? launch_sweep_t<M, R, false, true, true>(b, L, s) : launch_sweep_t<M, R, false, true, false>(b, L, s);
}

}  // namespace hp
