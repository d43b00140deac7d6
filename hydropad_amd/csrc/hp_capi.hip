// hp_capi.hip -- the extern "C" boundary (include/hydropad_gpu.h): device
// state ownership, host<->device copies and the step orchestration of
// hydro/ppm_lxlylz.F90 and hydropad/evolve_sys.F90 on one stream.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/hydropad_gpu.h"
#include "hp_core.h"
#include "hp_cosmo.h"
#include "hp_kernels.h"

static_assert(sizeof(hp_params) == sizeof(hp::Params), "hp_params and hp::Params must match");
static_assert(sizeof(hp_cosmo) == sizeof(hp::Cosmo), "hp_cosmo and hp::Cosmo must match");
static_assert(HP_FP64 == hp::PREC_FP64 && HP_MIXED == hp::PREC_MIXED && HP_FP64_EXACT == hp::PREC_FP64_EXACT,
              "precision codes");

namespace {

constexpr int NROT_ORDER[6][3] = {{1, 2, 3}, {3, 2, 1}, {2, 3, 1}, {1, 3, 2}, {3, 1, 2}, {2, 1, 3}};

#define HPCHK(expr)                          \
    do {                                     \
        if ((expr) != hipSuccess) return HP_E_HIP; \
    } while (0)

// Scalar integrate / dual energy on the device for host-side unit vectors.
__global__ void integrate_cells_kernel(hp::IScal S, int n, const double* __restrict__ f10, double* st5,
                                       const double* __restrict__ e3)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* f = f10 + 10 * (long)t;
    double* s = st5 + 5 * (long)t;
    const double* e = e3 + 3 * (long)t;
    hp::Flux F1, F2;
    F1.p = f[0]; F2.p = f[1]; F1.r = f[2]; F2.r = f[3]; F1.u = f[4]; F2.u = f[5];
    F1.v = f[6]; F2.v = f[7]; F1.w = f[8]; F2.w = f[9];
    double pres = s[0], rho = s[1], vx = s[2], vy = s[3], vz = s[4];
    // integrate() takes etot and eint from its caller (integrate.F90:1-2) and
    // always adds the gravity terms, zero when ghalf is (integrate.F90:43-44,56-57)
    hp::integrate_cell_e(S, F1, F2, pres, rho, vx, vy, vz, e[0], e[1], e[2], true);
    s[0] = pres; s[1] = rho; s[2] = vx; s[3] = vy; s[4] = vz;
}

__global__ void dualenergy_kernel(hp::IScal S, const double* __restrict__ in8, double* out2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double eintnew = in8[5], pnew = 0.0;
    hp::dual_energy(S, in8[0], in8[1], in8[2], in8[3], in8[4], in8[6], in8[7], eintnew, pnew);
    out2[0] = pnew;
    out2[1] = eintnew;
}

hp::IScal iscal_of(const hp_params& P)
{
    hp::IScal S;
    S.rdtath = P.rdtath; S.rdx = P.rdx; S.gamma = P.gamma; S.rgamma1 = P.rgamma1;
    S.eta1 = P.eta1; S.eta2 = P.eta2;
    return S;
}

}  // namespace

struct hp_ctx {
    int nx, ny, nz, nb;
    int precision;
    size_t ncell;
    void* pool;
    void* pool3;           // third field set, allocated on first use (expansion steps)
    double* f[3][5];       // field sets; sweeps read f[cur] and write another set
    int cur;
    int saved;             // set holding the kept start-of-step state (hp_savetn), or -1
    double* gforce;
    bool has_grav;
    hp::Params* d_prm;
    hp::DevStatus* d_st;
    hp_params h_prm;
    hipStream_t stream;
    int device;
    int seglen;
    int xcpw;
    int block;
    bool timing;
    double* ppos;          // particles (3,np): positions, velocities (the reference's ppos/pvel)
    double* pvel;
    long long np, pcap;
    double dtat;           // the kicks' time step (the reference's module scalar dtat)
    std::vector<hipEvent_t> ev_pool;     // recycled events
    std::vector<hipEvent_t> ev_used[3];  // start/end pairs per direction
};

static hipEvent_t take_event(hp_ctx* c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static bool is_cosmo(const hp_params& P) { return P.dat != 0.0 || P.at != 1.0 || P.rat != 1.0; }

// The set a sweep from f[cur] writes: neither cur nor the kept state.
static int out_set(const hp_ctx* c)
{
    for (int k = 0; k < 3; ++k)
        if (k != c->cur && k != c->saved) return k;
    return -1;
}

static bool ensure_third_set(hp_ctx* c)
{
    if (c->pool3) return true;
    const size_t align = 256;
    const size_t fstride = (c->ncell * sizeof(double) + align - 1) / align * align;
    if (hipMalloc(&c->pool3, 5 * fstride) != hipSuccess) { c->pool3 = nullptr; return false; }
    char* p = (char*)c->pool3;
    for (int f = 0; f < 5; ++f) { c->f[2][f] = (double*)p; p += fstride; }
    return true;
}

static hp::BoxView view_of(hp_ctx* c, int out)
{
    hp::BoxView b;
    for (int i = 0; i < 5; ++i) {
        b.in[i] = c->f[c->cur][i];
        b.out[i] = c->f[out][i];
    }
    b.gforce = c->has_grav ? c->gforce : nullptr;
    b.prm = c->d_prm;
    b.st = c->d_st;
    b.nx = c->nx; b.ny = c->ny; b.nz = c->nz; b.nb = c->nb;
    return b;
}

// One sweep f[cur] -> f[out_set]; with expand, the expansion step against the
// kept state f[saved] with the device factors st->fac is fused into it.
static int sweep(hp_ctx* c, int dir, int ghost_mask, int do_speed, bool expand = false)
{
    const int out = out_set(c);
    if (out < 0 || (out == 2 && !c->pool3)) return HP_E_STATE;
    if (expand && c->saved < 0) return HP_E_STATE;
    hp::SweepLaunch L;
    for (int f = 0; f < 4; ++f) L.old[f] = expand ? c->f[c->saved][f == 0 ? 0 : f + 1] : nullptr;
    L.fac = (const double*)((const char*)c->d_st + offsetof(hp::DevStatus, fac));
    L.dir = dir;
    L.ghost_mask = ghost_mask;
    L.do_speed = do_speed;
    L.seglen = c->seglen;
    L.xcpw = c->xcpw;
    L.block = c->block;
    L.precision = c->precision;
    L.grav = c->has_grav;
    L.cosmo = is_cosmo(c->h_prm);
    L.g53 = hp::params_g53(*(const hp::Params*)&c->h_prm);
    hp::BoxView b = view_of(c, out);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (!e0 || !e1 || hipEventRecord(e0, c->stream) != hipSuccess) return HP_E_HIP;
    }
    if (hp::launch_sweep(b, L, c->stream) != hipSuccess) return HP_E_HIP;
    if (c->timing) {
        if (hipEventRecord(e1, c->stream) != hipSuccess) return HP_E_HIP;
        c->ev_used[dir - 1].push_back(e0);
        c->ev_used[dir - 1].push_back(e1);
    }
    c->cur = out;
    return HP_OK;
}

// The reference stops the run when a Riemann Newton iteration does not
// converge (riemann.F90:60-65); here the device counter, cumulative since the
// last hp_reset_counters, turns into HP_E_NEWTON at the call's return.
static int newton_rc(hp_ctx* c)
{
    unsigned long long n = 0;
    HPCHK(hipMemcpyAsync(&n, &c->d_st->newton_nonconv, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    return n ? HP_E_NEWTON : HP_OK;
}

// The N-body part of a step (hydropad/evolve_sys.F90:63-84): kick, drift,
// kick over the particles with the context's gforce, their speeds folded into
// the CFL maximum (speed.F90:43-56).
static int particle_step(hp_ctx* c)
{
    if (c->np <= 0) return HP_OK;
    HPCHK(hp::launch_particles(3, c->ppos, c->pvel, c->np, c->has_grav ? c->gforce : nullptr, c->nx, c->ny, c->nz,
                               c->nb, c->d_prm, c->dtat, c->d_st, 1, c->stream));
    return HP_OK;
}

extern "C" {

int hp_params_init(double gamma, int ngrid_max, double dt, hp_params* o)
{
    if (!o || ngrid_max <= 0 || !(gamma > 1.0)) return HP_E_ARG;
    memset(o, 0, sizeof(*o));
    o->gamma = gamma;
    o->rgamma = 1.0 / gamma;
    o->gamma1 = (gamma + 1.0) / (gamma - 1.0);
    o->rgamma1 = 1.0 / (gamma - 1);
    o->m = (gamma - 1.0) / (2.0 * gamma);
    o->rm = 1.0 / o->m;
    o->gf = (gamma + 1.0) / (2.0 * gamma);
    o->dx = (double)(1.0f / (float)ngrid_max);   // FP32 in the reference (initialize_parameters.F90:23)
    o->rdx = 1.0 / o->dx;
    o->dt = dt;
    o->at = 1.0;
    o->dat = 0.0;
    o->rat = 1.0 / o->at;
    o->rdtath = dt / 1.0;
    o->flatvalue = 0.15;
    o->mingradflat = 0.25;
    o->eta1 = 1e-3;
    o->eta2 = 0.1;
    o->dmax = 2.0e10;
    return HP_OK;
}

int hp_create(int nx, int ny, int nz, int nbound, int precision, hp_ctx** out)
{
    if (!out) return HP_E_ARG;
    *out = nullptr;
    if (nbound != 3) return HP_E_ARG;   // the STENCIL path hard-wires a 6-cell stencil (ppmsolver.F90:82-84)
    if (nx < 2 * nbound + 3 || ny < 2 * nbound + 3 || nz < 2 * nbound + 3) return HP_E_ARG;
    if (precision != HP_FP64 && precision != HP_MIXED && precision != HP_FP64_EXACT) return HP_E_ARG;
    // the sweeps index cells of one box with 32-bit arithmetic: nx*ny*nz < 2^31
    // (a 1290^3 box is already past the 288 GB of HBM with two field sets)
    if ((long long)nx * ny * nz >= (1LL << 31)) return HP_E_ARG;
    hp_ctx* c = new (std::nothrow) hp_ctx();
    if (!c) return HP_E_NOMEM;
    c->nx = nx; c->ny = ny; c->nz = nz; c->nb = nbound;
    c->precision = precision;
    c->ncell = (size_t)nx * ny * nz;
    c->seglen = 0;
    c->xcpw = 0;
    c->block = 256;
    c->timing = false;
    c->ppos = c->pvel = nullptr;
    c->np = c->pcap = 0;
    c->dtat = 0.0;
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return HP_E_HIP; }
    const size_t fbytes = c->ncell * sizeof(double);
    const size_t align = 256;
    const size_t fstride = (fbytes + align - 1) / align * align;
    const size_t bytes = 13 * fstride + 2 * align + sizeof(hp::Params) + sizeof(hp::DevStatus);
    if (hipMalloc(&c->pool, bytes) != hipSuccess) { delete c; return HP_E_NOMEM; }
    char* p = (char*)c->pool;
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 5; ++f) { c->f[b][f] = (double*)p; p += fstride; }
    c->gforce = (double*)p; p += 3 * fstride;
    c->d_prm = (hp::Params*)p; p += align;
    c->d_st = (hp::DevStatus*)p;
    c->cur = 0;
    c->saved = -1;
    c->pool3 = nullptr;
    for (int f = 0; f < 5; ++f) c->f[2][f] = nullptr;
    c->has_grav = false;
    if (hipStreamCreate(&c->stream) != hipSuccess) { hipFree(c->pool); delete c; return HP_E_HIP; }
    // dx from the largest interior extent (initialize_parameters.F90:23)
    hp_params_init(5.0 / 3.0, std::max(nx, std::max(ny, nz)) - 2 * nbound, 0.0, &c->h_prm);
    hp::DevStatus z;
    memset(&z, 0, sizeof(z));
    if (hipMemcpyAsync(c->d_prm, &c->h_prm, sizeof(hp_params), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->d_st, &z, sizeof(z), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        hp_destroy(c);
        return HP_E_HIP;
    }
    *out = c;
    return HP_OK;
}

void hp_destroy(hp_ctx* c)
{
    if (!c) return;
    if (c->stream) { hipStreamSynchronize(c->stream); hipStreamDestroy(c->stream); }
    for (int d = 0; d < 3; ++d)
        for (hipEvent_t e : c->ev_used[d]) hipEventDestroy(e);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    if (c->pool) hipFree(c->pool);
    if (c->pool3) hipFree(c->pool3);
    if (c->ppos) hipFree(c->ppos);
    delete c;
}

int hp_set_precision(hp_ctx* c, int precision)
{
    if (!c || (precision != HP_FP64 && precision != HP_MIXED && precision != HP_FP64_EXACT)) return HP_E_ARG;
    c->precision = precision;
    return HP_OK;
}

static int upload_impl(hp_ctx* c, const double* const src[5], const double* gforce, hipMemcpyKind kind)
{
    if (!c) return HP_E_ARG;
    for (int f = 0; f < 5; ++f)
        if (!src[f]) return HP_E_ARG;
    const size_t bytes = c->ncell * sizeof(double);
    c->saved = -1;                               // an upload discards the kept state
    for (int f = 0; f < 5; ++f) HPCHK(hipMemcpyAsync(c->f[c->cur][f], src[f], bytes, kind, c->stream));
    if (gforce) HPCHK(hipMemcpyAsync(c->gforce, gforce, 3 * bytes, kind, c->stream));
    c->has_grav = gforce != nullptr;
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_upload(hp_ctx* c, const double* p, const double* rho, const double* vx, const double* vy, const double* vz,
              const double* gforce)
{
    const double* s[5] = {p, rho, vx, vy, vz};
    return upload_impl(c, s, gforce, hipMemcpyHostToDevice);
}

int hp_upload_gforce(hp_ctx* c, const double* gforce)
{
    if (!c) return HP_E_ARG;
    if (gforce)
        HPCHK(hipMemcpyAsync(c->gforce, gforce, 3 * c->ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->has_grav = gforce != nullptr;
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_upload_device(hp_ctx* c, const double* p, const double* rho, const double* vx, const double* vy,
                     const double* vz, const double* gforce)
{
    const double* s[5] = {p, rho, vx, vy, vz};
    return upload_impl(c, s, gforce, hipMemcpyDeviceToDevice);
}

static int download_impl(hp_ctx* c, double* const dst[5], hipMemcpyKind kind)
{
    if (!c) return HP_E_ARG;
    for (int f = 0; f < 5; ++f)
        if (!dst[f]) return HP_E_ARG;
    // complete the ghost layer (faces, edges, corners) as exchange_mesh would
    HPCHK(hp::launch_fill(c->f[c->cur], c->nx, c->ny, c->nz, c->nb, 0, c->stream));
    const size_t bytes = c->ncell * sizeof(double);
    for (int f = 0; f < 5; ++f) HPCHK(hipMemcpyAsync(dst[f], c->f[c->cur][f], bytes, kind, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_download(hp_ctx* c, double* p, double* rho, double* vx, double* vy, double* vz)
{
    double* d[5] = {p, rho, vx, vy, vz};
    return download_impl(c, d, hipMemcpyDeviceToHost);
}

int hp_download_device(hp_ctx* c, double* p, double* rho, double* vx, double* vy, double* vz)
{
    double* d[5] = {p, rho, vx, vy, vz};
    return download_impl(c, d, hipMemcpyDeviceToDevice);
}

int hp_set_params(hp_ctx* c, const hp_params* prm)
{
    if (!c || !prm) return HP_E_ARG;
    if (!(prm->dx > 0.0) || !(prm->gamma > 1.0)) return HP_E_ARG;
    c->h_prm = *prm;
    HPCHK(hipMemcpyAsync(c->d_prm, &c->h_prm, sizeof(hp_params), hipMemcpyHostToDevice, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_ppmsolver(hp_ctx* c, int ndirection)
{
    if (!c || ndirection < 1 || ndirection > 3) return HP_E_ARG;
    const int r = sweep(c, ndirection, 7, 0);
    return r ? r : newton_rc(c);
}

int hp_exchange_mesh(hp_ctx* c, int direction)
{
    if (!c || direction < 0 || direction > 3) return HP_E_ARG;
    HPCHK(hp::launch_fill(c->f[c->cur], c->nx, c->ny, c->nz, c->nb, direction, c->stream));
    return HP_OK;
}

int hp_ppm_lxlylz(hp_ctx* c, int nrot)
{
    if (!c || nrot < 1 || nrot > 6) return HP_E_ARG;
    for (int s = 0; s < 3; ++s) {
        int r = sweep(c, NROT_ORDER[nrot - 1][s], 7, 0);
        if (r) return r;
    }
    return newton_rc(c);
}

int hp_speed(hp_ctx* c, double* velmax)
{
    if (!c || !velmax) return HP_E_ARG;
    HPCHK(hp::launch_zero_status(c->d_st, 1, c->stream));
    const double* f[5] = {c->f[c->cur][0], c->f[c->cur][1], c->f[c->cur][2], c->f[c->cur][3], c->f[c->cur][4]};
    HPCHK(hp::launch_speed(f, c->nx, c->ny, c->nz, c->nb, c->d_prm, c->d_st, c->stream));
    HPCHK(hp::launch_particle_vmax(c->pvel, c->np, c->d_st, c->stream));   // speed.F90:43-56
    unsigned long long bits = 0;
    HPCHK(hipMemcpyAsync(&bits, &c->d_st->velmax_bits, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HPCHK(hp::launch_zero_status(c->d_st, 1, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    memcpy(velmax, &bits, sizeof(double));
    return HP_OK;
}

int hp_evolve(hp_ctx* c, int first_step, int nsteps, double cour, double tstop, double* t, double* dt)
{
    if (!c || !t || !dt || nsteps < 0 || first_step < 1 || !(cour > 0.0)) return HP_E_ARG;
    if (is_cosmo(c->h_prm)) return HP_E_STATE;   // the cosmological schedule is driven per step by the caller
    c->h_prm.dt = *dt;
    c->h_prm.rdtath = *dt / 1.0;
    HPCHK(hipMemcpyAsync(c->d_prm, &c->h_prm, sizeof(hp_params), hipMemcpyHostToDevice, c->stream));
    double tt = *t;
    HPCHK(hipMemcpyAsync(&c->d_st->t, &tt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCHK(hp::launch_zero_status(c->d_st, 1, c->stream));
    for (int nstep = first_step; nstep < first_step + nsteps; ++nstep) {
        const int nrot = (nstep == 1) ? 1 : nstep % 6 + 1;          // evolve_sys.F90:30-31
        for (int s = 0; s < 3; ++s) {
            int r = sweep(c, NROT_ORDER[nrot - 1][s], 7, s == 2 ? 1 : 0);
            if (r) return r;
        }
        if (int r = particle_step(c)) return r;
        HPCHK(hp::launch_timedet(c->d_prm, c->d_st, cour, tstop, c->stream));
    }
    HPCHK(hipMemcpyAsync(&c->h_prm, c->d_prm, sizeof(hp_params), hipMemcpyDeviceToHost, c->stream));
    HPCHK(hipMemcpyAsync(&tt, &c->d_st->t, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    *t = tt;
    *dt = c->h_prm.dt;
    return newton_rc(c);
}

// ---- comoving boxes ------------------------------------------------------------

int hp_cosmo_init(double a_start, double cour, double ca, double tstop, hp_cosmo* s)
{
    if (!s || !(a_start > 0.0) || !(cour > 0.0) || !(ca > 0.0)) return HP_E_ARG;
    hp::Cosmo c;
    hp::cosmo_init(a_start, cour, ca, tstop, c);
    memcpy(s, &c, sizeof(c));
    return HP_OK;
}

int hp_timedet_cosmo(hp_cosmo* s, double velmax, double dx)
{
    if (!s || !(dx > 0.0) || !(velmax >= 0.0)) return HP_E_ARG;
    hp::Cosmo c;
    memcpy(&c, s, sizeof(c));
    hp::cosmo_timedet(c, velmax, dx);
    memcpy(s, &c, sizeof(c));
    return HP_OK;
}

int hp_cosmo_params(const hp_cosmo* s, hp_params* prm)
{
    if (!s || !prm) return HP_E_ARG;
    prm->at = s->at;
    prm->dat = s->dat;
    prm->rat = s->rat;
    prm->dt = s->dt;
    prm->rdtath = s->rdtath;
    return HP_OK;
}

int hp_savetn(hp_ctx* c)
{
    if (!c) return HP_E_ARG;
    if (!ensure_third_set(c)) return HP_E_NOMEM;
    c->saved = c->cur;            // sweeps now ping-pong between the other two sets
    return HP_OK;
}

int hp_expand(hp_ctx* c, double afact)
{
    if (!c) return HP_E_ARG;
    if (c->saved < 0 || c->saved == c->cur) return HP_E_STATE;
    const double* old[4] = {c->f[c->saved][0], c->f[c->saved][2], c->f[c->saved][3], c->f[c->saved][4]};
    HPCHK(hp::launch_expand(c->f[c->cur], old, (long)c->ncell, afact, c->stream));
    return HP_OK;
}

int hp_evolve_cosmo(hp_ctx* c, int first_step, int nsteps, hp_cosmo* s)
{
    if (!c || !s || nsteps < 0 || first_step < 1) return HP_E_ARG;
    if (!(s->at > 0.0) || !(s->ath > 0.0) || !(s->cour > 0.0)) return HP_E_ARG;
    if (!ensure_third_set(c)) return HP_E_NOMEM;
    hp::Cosmo cs;
    memcpy(&cs, s, sizeof(cs));
    hp_cosmo_params(s, &c->h_prm);
    double fac[3];
    hp::cosmo_factors(cs, fac);
    HPCHK(hipMemcpyAsync(c->d_prm, &c->h_prm, sizeof(hp_params), hipMemcpyHostToDevice, c->stream));
    HPCHK(hipMemcpyAsync(&c->d_st->cosmo, &cs, sizeof(cs), hipMemcpyHostToDevice, c->stream));
    HPCHK(hipMemcpyAsync(&c->d_st->fac[0], fac, sizeof(fac), hipMemcpyHostToDevice, c->stream));
    HPCHK(hp::launch_zero_status(c->d_st, 1, c->stream));
    for (int nstep = first_step; nstep < first_step + nsteps; ++nstep) {
        const int nrot = (nstep == 1) ? 1 : nstep % 6 + 1;          // evolve_sys.F90:30-31
        c->saved = c->cur;                                         // start-of-step state
        for (int k = 0; k < 3; ++k) {
            int r = sweep(c, NROT_ORDER[nrot - 1][k], 7, k == 2 ? 1 : 0, k == 2);
            if (r) return r;
        }
        if (int r = particle_step(c)) return r;
        HPCHK(hp::launch_timedet_cosmo(c->d_prm, c->d_st, c->stream));
    }
    HPCHK(hipMemcpyAsync(&c->h_prm, c->d_prm, sizeof(hp_params), hipMemcpyDeviceToHost, c->stream));
    HPCHK(hipMemcpyAsync(&cs, &c->d_st->cosmo, sizeof(cs), hipMemcpyDeviceToHost, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    memcpy(s, &cs, sizeof(cs));
    return newton_rc(c);
}

int hp_set_tiling(hp_ctx* c, int x_chunks, int yz_seglen, int block)
{
    if (!c || x_chunks < 0 || yz_seglen < 0 || block < 0 || block > 256 || block % 64 != 0) return HP_E_ARG;
    c->xcpw = x_chunks;
    c->seglen = yz_seglen;
    c->block = block > 0 ? block : 256;
    return HP_OK;
}

int hp_sweep_timing_enable(hp_ctx* c, int on)
{
    if (!c) return HP_E_ARG;
    HPCHK(hipStreamSynchronize(c->stream));
    for (int d = 0; d < 3; ++d) {
        for (hipEvent_t e : c->ev_used[d]) c->ev_pool.push_back(e);
        c->ev_used[d].clear();
    }
    c->timing = on != 0;
    return HP_OK;
}

int hp_sweep_timing(hp_ctx* c, double ms[3], int64_t launches[3])
{
    if (!c || !ms || !launches) return HP_E_ARG;
    HPCHK(hipStreamSynchronize(c->stream));
    for (int d = 0; d < 3; ++d) {
        double tot = 0.0;
        const std::vector<hipEvent_t>& v = c->ev_used[d];
        for (size_t i = 0; i + 1 < v.size(); i += 2) {
            float m = 0.f;
            HPCHK(hipEventElapsedTime(&m, v[i], v[i + 1]));
            tot += m;
        }
        ms[d] = tot;
        launches[d] = (int64_t)(v.size() / 2);
    }
    return HP_OK;
}

int hp_status(hp_ctx* c, hp_counters* out)
{
    if (!c || !out) return HP_E_ARG;
    hp::DevStatus s;
    HPCHK(hipMemcpyAsync(&s, c->d_st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof(*out));
    out->newton_iters_max = s.newton_max;
    out->newton_nonconv = (int64_t)s.newton_nonconv;
    out->neg_state_clamps = (int64_t)s.clamps;
    out->particles_outside = (int64_t)s.outside;
    return s.newton_nonconv ? HP_E_NEWTON : HP_OK;
}

// ---- particles (nbody/) ------------------------------------------------------

int hp_particles_upload(hp_ctx* c, int64_t np, const double* ppos, const double* pvel, double dtat)
{
    if (!c || np < 0 || (np > 0 && (!ppos || !pvel))) return HP_E_ARG;
    if (np >= (1LL << 40)) return HP_E_ARG;
    HPCHK(hipStreamSynchronize(c->stream));
    if (np > c->pcap) {
        if (c->ppos) hipFree(c->ppos);
        c->ppos = c->pvel = nullptr;
        c->np = c->pcap = 0;
        const size_t half = ((size_t)np * 3 * sizeof(double) + 255) / 256 * 256;
        void* m = nullptr;
        if (hipMalloc(&m, 2 * half) != hipSuccess) return HP_E_NOMEM;
        c->ppos = (double*)m;
        c->pvel = (double*)((char*)m + half);
        c->pcap = np;
    }
    const size_t bytes = (size_t)np * 3 * sizeof(double);
    if (np > 0) {
        HPCHK(hipMemcpyAsync(c->ppos, ppos, bytes, hipMemcpyHostToDevice, c->stream));
        HPCHK(hipMemcpyAsync(c->pvel, pvel, bytes, hipMemcpyHostToDevice, c->stream));
    }
    c->np = np;
    c->dtat = dtat;
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_particles_download(hp_ctx* c, double* ppos, double* pvel)
{
    if (!c || (c->np > 0 && (!ppos || !pvel))) return HP_E_ARG;
    const size_t bytes = (size_t)c->np * 3 * sizeof(double);
    if (c->np > 0) {
        HPCHK(hipMemcpyAsync(ppos, c->ppos, bytes, hipMemcpyDeviceToHost, c->stream));
        HPCHK(hipMemcpyAsync(pvel, c->pvel, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_kick(hp_ctx* c, double dtat)
{
    if (!c) return HP_E_ARG;
    HPCHK(hp::launch_particles(1, c->ppos, c->pvel, c->np, c->has_grav ? c->gforce : nullptr, c->nx, c->ny, c->nz,
                               c->nb, c->d_prm, dtat, c->d_st, 0, c->stream));
    return HP_OK;
}

int hp_drift(hp_ctx* c)
{
    if (!c) return HP_E_ARG;
    HPCHK(hp::launch_particles(2, c->ppos, c->pvel, c->np, nullptr, c->nx, c->ny, c->nz, c->nb, c->d_prm, 0.0, c->d_st,
                               0, c->stream));
    return HP_OK;
}

int hp_set_newton_limit(hp_ctx* c, int nmax)
{
    if (!c || nmax < 0) return HP_E_ARG;
    HPCHK(hipMemcpyAsync(&c->d_st->newton_cap, &nmax, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_reset_counters(hp_ctx* c)
{
    if (!c) return HP_E_ARG;
    HPCHK(hp::launch_zero_status(c->d_st, 2, c->stream));
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_sync(hp_ctx* c)
{
    if (!c) return HP_E_ARG;
    HPCHK(hipStreamSynchronize(c->stream));
    return HP_OK;
}

int hp_integrate_cell(const hp_params* prm, const double* pm1, const double* pm2, const double* rhom1,
                      const double* rhom2, const double* vxm1, const double* vxm2, const double* vym1,
                      const double* vym2, const double* vzm1, const double* vzm2, double* pres, double* rho,
                      double* vx, double* vy, double* vz, const double* etot, const double* eint,
                      const double* ghalf)
{
    if (!prm || !pm1 || !pm2 || !rhom1 || !rhom2 || !vxm1 || !vxm2 || !vym1 || !vym2 || !vzm1 || !vzm2 || !pres ||
        !rho || !vx || !vy || !vz || !etot || !eint || !ghalf)
        return HP_E_ARG;
    double h[18] = {*pm1, *pm2, *rhom1, *rhom2, *vxm1, *vxm2, *vym1, *vym2, *vzm1, *vzm2,
                    *pres, *rho, *vx, *vy, *vz, *etot, *eint, *ghalf};
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(h)) != hipSuccess) return HP_E_NOMEM;
    int rc = HP_OK;
    if (hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) rc = HP_E_HIP;
    if (rc == HP_OK) {
        hipLaunchKernelGGL(integrate_cells_kernel, dim3(1), dim3(64), 0, 0, iscal_of(*prm), 1, d, d + 10, d + 15);
        if (hipGetLastError() != hipSuccess || hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
            rc = HP_E_HIP;
    }
    hipFree(d);
    if (rc == HP_OK) { *pres = h[10]; *rho = h[11]; *vx = h[12]; *vy = h[13]; *vz = h[14]; }
    return rc;
}

int hp_dualenergy_cell(const hp_params* prm, double etotnew, double rhonew, double vx, double vy, double vz,
                       double eintnew_in, double eint_old, double pres_old, double* pnew, double* eintnew)
{
    if (!prm || !pnew || !eintnew) return HP_E_ARG;
    double h[10] = {etotnew, rhonew, vx, vy, vz, eintnew_in, eint_old, pres_old, 0.0, 0.0};
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(h)) != hipSuccess) return HP_E_NOMEM;
    int rc = HP_OK;
    if (hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) rc = HP_E_HIP;
    if (rc == HP_OK) {
        hipLaunchKernelGGL(dualenergy_kernel, dim3(1), dim3(64), 0, 0, iscal_of(*prm), d, d + 8);
        if (hipGetLastError() != hipSuccess || hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
            rc = HP_E_HIP;
    }
    hipFree(d);
    if (rc == HP_OK) { *pnew = h[8]; *eintnew = h[9]; }
    return rc;
}

}  // extern "C"
