// hp_batch.hip -- a batch of independent periodic boxes resident on one GPU,
// advanced together: per step a plan kernel builds the work tables from each
// box's own step number (its own nrot order, hydropad/evolve_sys.F90:30-31),
// persistent sweep kernels run the three sweep slots, and one timedet kernel
// gives every box its own next dt (hydropad/functions.F90:2-35, or the
// comoving schedule of :38-151).  The step's launch sequence is fixed for a
// given mix of precisions, so it is captured once in a hipGraph and
// replayed.  Boxes can be added and removed between steps (continuous
// batching); memory is admitted against an HBM budget.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/hydropad_gpu.h"
#include "hp_batch.h"
#include "hp_core.h"
#include "hp_cosmo.h"
#include "hp_kernels.h"

static_assert(sizeof(hp_box_status) % 8 == 0, "hp_box_status layout");

namespace {

#define HPCHK(expr)                                \
    do {                                           \
        if ((expr) != hipSuccess) return HP_E_HIP; \
    } while (0)

constexpr int ORDER[6][3] = {{1, 2, 3}, {3, 2, 1}, {2, 3, 1}, {1, 3, 2}, {3, 1, 2}, {2, 1, 3}};

// One thread per plan table: the boxes whose sweep `slot` of their next step
// is (dir, precision, fused expansion), with prefix sums of their work units.
__global__ void plan_kernel(const hp::BatchBox* __restrict__ boxes, hp::BatchPlan* plan)
{
    const int key = threadIdx.x;
    if (key >= hp::BATCH_NTABLE) return;
    const int g53 = key & 1, exp = (key >> 1) & 1, prec = (key >> 2) & 1, sd = key >> 3;
    const int slot = sd / 3, dir = sd % 3 + 1;
    hp::BatchTable& T = plan->t[key];
    const int nbox = plan->nbox;
    int n = 0, total = 0;
    for (int b = 0; b < nbox; ++b) {
        const hp::BatchBox& X = boxes[b];
        if (!X.active || X.st->done) continue;
        const long long step = X.st->nstep_done + 1;
        const int nrot = (step == 1) ? 1 : (int)(step % 6) + 1;
        const int d = ORDER[nrot - 1][slot];
        const int e = (slot == 2 && X.nsets == 3) ? 1 : 0;
        if (d != dir || X.prec != prec || e != exp || X.g53 != g53) continue;
        T.box[n] = b;
        T.start[n] = total;
        total += X.units[dir - 1];
        ++n;
    }
    T.n = n;
    T.total = total;
}

// Per box: consume the step's CFL maximum, set the next step's scalars,
// count the step.
__global__ void batch_timedet_kernel(const hp::BatchBox* __restrict__ boxes, const hp::BatchPlan* __restrict__ plan)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= plan->nbox) return;
    const hp::BatchBox& X = boxes[b];
    if (!X.active) return;
    hp::DevStatus* st = X.st;
    if (st->done) return;
    hp::Params* prm = X.prm;
    const double velmax = __longlong_as_double((long long)st->velmax_bits);
    if (st->comoving) {
        hp::Cosmo c = st->cosmo;
        hp::cosmo_timedet(c, velmax, prm->dx);
        st->cosmo = c;
        prm->at = c.at; prm->dat = c.dat; prm->rat = c.rat; prm->dt = c.dt; prm->rdtath = c.rdtath;
        hp::cosmo_factors(c, st->fac);
    } else {                                             // functions.F90:2-35
        const double told = st->t;
        const double rdtc = velmax / (st->cour * prm->dx);
        double dt = (velmax == 0.0) ? 0.1 : 1.0 / rdtc;
        double t = told + dt;
        if (t > st->tstop) { t = st->tstop; dt = t - told; }
        st->t = t;
        prm->dt = dt;
        prm->rdtath = dt / 1.0;
    }
    st->velmax_bits = 0ull;
    st->nstep_done += 1;
    if (st->nstep_done >= st->nstep_target) st->done = 1;
}

// dt of the first step of a static box from its initial CFL maximum
// (the same timedet with t = 0).
__global__ void first_dt_kernel(hp::Params* prm, hp::DevStatus* st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double velmax = __longlong_as_double((long long)st->velmax_bits);
    const double rdtc = velmax / (st->cour * prm->dx);
    double dt = (velmax == 0.0) ? 0.1 : 1.0 / rdtc;
    double t = dt;
    if (t > st->tstop) { t = st->tstop; dt = t; }
    st->t = t;
    prm->dt = dt;
    prm->rdtath = dt / 1.0;
    st->velmax_bits = 0ull;
}

__host__ __device__ inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Seeded planar Sod tube written on the device, ghosts included (periodic images).
__global__ void __launch_bounds__(256) sod_ic_kernel(double* p, double* rho, double* vx, double* vy, double* vz,
                                                     int n, int nb, int axis, double x0, int flip)
{
    const int na = n + 2 * nb;
    const long tot = (long)na * na * na;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (long)gridDim.x * blockDim.x) {
        const long i = q % na, r = q / na;
        const long j = r % na, k = r / na;
        const long c = axis == 0 ? i : (axis == 1 ? j : k);
        const long ia = ((c - nb) % n + n) % n;
        const double x = ((double)ia + 0.5) / (double)n;
        bool left = x < x0;
        if (flip) left = !left;
        rho[q] = left ? 1.0 : 0.125;
        p[q] = left ? 1.0 : 0.1;
        vx[q] = 0.0; vy[q] = 0.0; vz[q] = 0.0;
    }
}

}  // namespace

struct hp_batch {
    int maxb;
    size_t budget, used;
    hipStream_t stream;
    int device, ncu;
    hp::BatchBox* d_boxes;
    hp::BatchPlan* d_plan;
    hp::Params* d_prm;
    hp::DevStatus* d_st;
    std::vector<hp::BatchBox> h_boxes;
    std::vector<void*> pools;
    std::vector<size_t> bytes;
    int nbox;
    hp_batch_tuning tune;
    bool use_graph;
    hipGraphExec_t gexec;
    unsigned gcombos;
    int ggrid;
    std::vector<hp_box_status> h_status;
    // pinned status copies for batch_status_async / batch_status_wait
    hp::DevStatus* pin_st[2];
    hp::Params* pin_prm[2];
    hipEvent_t pin_ev[2];
    int pin_n[2];
};

namespace {

size_t field_bytes(int nx, int ny, int nz)
{
    const size_t align = 256;
    return ((size_t)nx * ny * nz * sizeof(double) + align - 1) / align * align;
}

int persistent_grid(const hp_batch* b)
{
    const int wpc = b->tune.waves_per_cu > 0 ? b->tune.waves_per_cu : 8;
    return std::max(1, b->ncu * wpc / 4);
}

void set_units(hp::BatchBox& X, const hp_batch_tuning& t)
{
    const int Nx = X.nx - 2 * X.nb, Ny = X.ny - 2 * X.nb, Nz = X.nz - 2 * X.nb;
    const long total = (long)Ny * Nz * X.nx;
    const long nchunks = (total + 57) / 58;
    X.xcpu = t.x_chunks > 0 ? t.x_chunks : 4;
    X.units[0] = (int)((nchunks + X.xcpu - 1) / X.xcpu);
    const int nsegq = t.yz_segments > 0 ? t.yz_segments : 1;
    for (int d = 0; d < 2; ++d) {
        const int Nax = d == 0 ? Ny : Nz;
        const int Nbb = d == 0 ? Nz : Ny;
        int seglen = (Nax + nsegq - 1) / nsegq;
        if (seglen < 1) seglen = 1;
        const int nseg = (Nax + seglen - 1) / seglen;
        X.seglen[d] = seglen;
        X.nseg[d] = nseg;
        const long items = (long)Nx * Nbb * nseg;
        X.units[d + 1] = (int)((items + 63) / 64);
    }
}

// bit (prec*4 + comoving*2 + g53) for every active box
unsigned combos_of(const hp_batch* b)
{
    unsigned m = 0;
    for (int i = 0; i < b->nbox; ++i) {
        const hp::BatchBox& X = b->h_boxes[i];
        if (X.active) m |= 1u << (X.prec * 4 + (X.nsets == 3 ? 2 : 0) + X.g53);
    }
    return m;
}

int push_box(hp_batch* b, int id)
{
    HPCHK(hipMemcpyAsync(b->d_boxes + id, &b->h_boxes[id], sizeof(hp::BatchBox), hipMemcpyHostToDevice, b->stream));
    return HP_OK;
}

int push_nbox(hp_batch* b)
{
    int n = 0;
    for (int i = 0; i < b->maxb; ++i)
        if (b->h_boxes[i].active) n = i + 1;
    b->nbox = n;
    HPCHK(hipMemcpyAsync(&b->d_plan->nbox, &b->nbox, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HPCHK(hipStreamSynchronize(b->stream));
    return HP_OK;
}

int enqueue_step(hp_batch* b, unsigned combos, int grid)
{
    hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(128), 0, b->stream, b->d_boxes, b->d_plan);
    for (int slot = 0; slot < 3; ++slot) {
        for (int prec = 0; prec < 2; ++prec) {
            for (int g = 0; g < 2; ++g) {
                const bool stat = combos & (1u << (prec * 4 + g)), como = combos & (1u << (prec * 4 + 2 + g));
                if (!stat && !como) continue;
                for (int dir = 1; dir <= 3; ++dir) {
                    if (slot < 2 || stat)
                        if (hp::launch_batch_sweep(b->d_boxes, b->d_plan, slot, dir, prec, 0, g, grid, b->stream) !=
                            hipSuccess)
                            return HP_E_HIP;
                    if (slot == 2 && como)
                        if (hp::launch_batch_sweep(b->d_boxes, b->d_plan, slot, dir, prec, 1, g, grid, b->stream) !=
                            hipSuccess)
                            return HP_E_HIP;
                }
            }
        }
    }
    const int tb = (b->maxb + 63) / 64;
    hipLaunchKernelGGL(batch_timedet_kernel, dim3(tb), dim3(64), 0, b->stream, b->d_boxes, b->d_plan);
    return hipGetLastError() == hipSuccess ? HP_OK : HP_E_HIP;
}

void drop_graph(hp_batch* b)
{
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->gcombos = 0;
    b->ggrid = 0;
}

int capture_step(hp_batch* b, unsigned combos, int grid)
{
    drop_graph(b);
    hipGraph_t g = nullptr;
    HPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int r = enqueue_step(b, combos, grid);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (r != HP_OK || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return HP_E_HIP;
    }
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->gexec = nullptr;
        return HP_E_HIP;
    }
    b->gcombos = combos;
    b->ggrid = grid;
    return HP_OK;
}

int find_slot(const hp_batch* b)
{
    for (int i = 0; i < b->maxb; ++i)
        if (!b->h_boxes[i].active) return i;
    return -1;
}

// Allocate a box's field sets, fill in its descriptor (not yet active).
int alloc_box(hp_batch* b, int id, int nx, int ny, int nz, int nb, int precision, int comoving)
{
    const size_t need = hp_batch_box_bytes(nx, ny, nz, comoving);
    if (b->used + need > b->budget) return HP_E_NOMEM;
    void* pool = nullptr;
    if (hipMalloc(&pool, need) != hipSuccess) return HP_E_NOMEM;
    hp::BatchBox X;
    memset(&X, 0, sizeof(X));
    X.nsets = comoving ? 3 : 2;
    const size_t fb = field_bytes(nx, ny, nz);
    char* p = (char*)pool;
    for (int s = 0; s < X.nsets; ++s)
        for (int f = 0; f < 5; ++f) { X.f[s][f] = (double*)p; p += fb; }
    X.prm = b->d_prm + id;
    X.st = b->d_st + id;
    X.nx = nx; X.ny = ny; X.nz = nz; X.nb = nb;
    X.prec = precision == HP_MIXED ? hp::PREC_MIXED : hp::PREC_FP64;
    X.active = 0;
    set_units(X, b->tune);
    b->h_boxes[id] = X;
    b->pools[id] = pool;
    b->bytes[id] = need;
    b->used += need;
    return HP_OK;
}

void free_box(hp_batch* b, int id)
{
    if (b->pools[id]) hipFree(b->pools[id]);
    b->pools[id] = nullptr;
    b->used -= b->bytes[id];
    b->bytes[id] = 0;
    memset(&b->h_boxes[id], 0, sizeof(hp::BatchBox));
}

// Device status + scalars of a new box, then activate it.
int start_box(hp_batch* b, int id, double gamma, double cour, double tstop, const hp_cosmo* cosmo, int nsteps)
{
    hp::BatchBox& X = b->h_boxes[id];
    hp_params P;
    const int ngrid = std::max(X.nx, std::max(X.ny, X.nz)) - 2 * X.nb;
    hp_params_init(gamma, ngrid, 0.0, &P);
    hp::DevStatus st;
    memset(&st, 0, sizeof(st));
    st.nstep_target = nsteps;
    st.cour = cour;
    st.tstop = tstop;
    st.comoving = cosmo != nullptr;
    if (cosmo) {
        hp_cosmo_params(cosmo, &P);
        memcpy(&st.cosmo, cosmo, sizeof(hp::Cosmo));
        hp::cosmo_factors(st.cosmo, st.fac);
    }
    HPCHK(hipMemcpyAsync(X.prm, &P, sizeof(P), hipMemcpyHostToDevice, b->stream));
    HPCHK(hipMemcpyAsync(X.st, &st, sizeof(st), hipMemcpyHostToDevice, b->stream));
    if (!cosmo) {
        const double* f[5] = {X.f[0][0], X.f[0][1], X.f[0][2], X.f[0][3], X.f[0][4]};
        HPCHK(hp::launch_speed(f, X.nx, X.ny, X.nz, X.nb, X.prm, X.st, b->stream));
        hipLaunchKernelGGL(first_dt_kernel, dim3(1), dim3(64), 0, b->stream, X.prm, X.st);
        HPCHK(hipGetLastError());
    }
    X.g53 = hp::params_g53(*(const hp::Params*)&P) ? 1 : 0;
    X.active = 1;
    if (push_box(b, id)) return HP_E_HIP;
    HPCHK(hipStreamSynchronize(b->stream));       // host state above lives on this stack
    return push_nbox(b);
}

}  // namespace

extern "C" {

size_t hp_batch_box_bytes(int nx, int ny, int nz, int comoving)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    return (size_t)(comoving ? 3 : 2) * 5 * field_bytes(nx, ny, nz);
}

int hp_batch_create(int max_boxes, size_t hbm_budget_bytes, hp_batch** out)
{
    if (!out || max_boxes < 1 || max_boxes > hp::BATCH_MAXB) return HP_E_ARG;
    *out = nullptr;
    hp_batch* b = new (std::nothrow) hp_batch();
    if (!b) return HP_E_NOMEM;
    b->maxb = max_boxes;
    b->budget = hbm_budget_bytes;
    b->used = 0;
    b->h_boxes.assign(max_boxes, hp::BatchBox());
    for (auto& X : b->h_boxes) memset(&X, 0, sizeof(X));
    b->pools.assign(max_boxes, nullptr);
    b->bytes.assign(max_boxes, 0);
    b->h_status.assign(max_boxes, hp_box_status());
    memset(&b->tune, 0, sizeof(b->tune));
    hipDeviceProp_t prop;
    if (hipGetDevice(&b->device) != hipSuccess || hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
        delete b;
        return HP_E_HIP;
    }
    b->ncu = prop.multiProcessorCount;
    if (hipStreamCreate(&b->stream) != hipSuccess) { delete b; return HP_E_HIP; }
    const size_t sz = sizeof(hp::BatchBox) * max_boxes + sizeof(hp::BatchPlan) + sizeof(hp::Params) * max_boxes +
                      sizeof(hp::DevStatus) * max_boxes + 4096;
    void* mem = nullptr;
    if (hipMalloc(&mem, sz) != hipSuccess) { hipStreamDestroy(b->stream); delete b; return HP_E_NOMEM; }
    char* p = (char*)mem;
    b->d_plan = (hp::BatchPlan*)p; p += (sizeof(hp::BatchPlan) + 255) / 256 * 256;
    b->d_boxes = (hp::BatchBox*)p; p += (sizeof(hp::BatchBox) * max_boxes + 255) / 256 * 256;
    b->d_prm = (hp::Params*)p; p += (sizeof(hp::Params) * max_boxes + 255) / 256 * 256;
    b->d_st = (hp::DevStatus*)p;
    std::vector<char> zero(sz, 0);
    if (hipMemcpy(mem, zero.data(), sz, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(mem);
        hipStreamDestroy(b->stream);
        delete b;
        return HP_E_HIP;
    }
    for (int k = 0; k < 2; ++k) {
        b->pin_st[k] = nullptr;
        b->pin_prm[k] = nullptr;
        b->pin_ev[k] = nullptr;
        b->pin_n[k] = 0;
    }
    b->nbox = 0;
    b->use_graph = false;
    b->gexec = nullptr;
    b->gcombos = 0;
    b->ggrid = 0;
    *out = b;
    return HP_OK;
}

void hp_batch_destroy(hp_batch* b)
{
    if (!b) return;
    hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (int k = 0; k < 2; ++k) {
        if (b->pin_st[k]) hipHostFree(b->pin_st[k]);
        if (b->pin_prm[k]) hipHostFree(b->pin_prm[k]);
        if (b->pin_ev[k]) hipEventDestroy(b->pin_ev[k]);
    }
    for (int i = 0; i < b->maxb; ++i)
        if (b->pools[i]) hipFree(b->pools[i]);
    hipFree(b->d_plan);                   // start of the descriptor block
    hipStreamDestroy(b->stream);
    delete b;
}

int hp_batch_add(hp_batch* b, int nx, int ny, int nz, int nbound, int precision, double gamma, double cour,
                 double tstop, const hp_cosmo* cosmo, int nsteps, const double* p, const double* rho,
                 const double* vx, const double* vy, const double* vz, int* id)
{
    if (!b || !id || !p || !rho || !vx || !vy || !vz || nsteps < 1 || !(cour > 0.0)) return HP_E_ARG;
    if (nbound != 3 || nx < 2 * nbound + 3 || ny < 2 * nbound + 3 || nz < 2 * nbound + 3) return HP_E_ARG;
    if ((long long)nx * ny * nz >= (1LL << 31)) return HP_E_ARG;   // 32-bit cell indexing in the sweeps
    if (precision != HP_FP64 && precision != HP_MIXED) return HP_E_ARG;
    if (!(gamma > 1.0)) return HP_E_ARG;
    const int slot = find_slot(b);
    if (slot < 0) return HP_E_NOMEM;
    int r = alloc_box(b, slot, nx, ny, nz, nbound, precision, cosmo != nullptr);
    if (r) return r;
    const hp::BatchBox& X = b->h_boxes[slot];
    const size_t bytes = (size_t)nx * ny * nz * sizeof(double);
    const double* src[5] = {p, rho, vx, vy, vz};
    for (int f = 0; f < 5; ++f)
        if (hipMemcpyAsync(X.f[0][f], src[f], bytes, hipMemcpyHostToDevice, b->stream) != hipSuccess) {
            free_box(b, slot);
            return HP_E_HIP;
        }
    r = start_box(b, slot, gamma, cour, tstop, cosmo, nsteps);
    if (r) { free_box(b, slot); push_nbox(b); return r; }
    *id = slot;
    return HP_OK;
}

int hp_batch_add_sod(hp_batch* b, int n, int precision, uint64_t seed, double cour, int nsteps, int* id)
{
    if (!b || !id || n < 3 || nsteps < 1 || !(cour > 0.0)) return HP_E_ARG;
    if ((long long)(n + 6) * (n + 6) * (n + 6) >= (1LL << 31)) return HP_E_ARG;
    if (precision != HP_FP64 && precision != HP_MIXED) return HP_E_ARG;
    const int nb = 3, na = n + 2 * nb;
    const int slot = find_slot(b);
    if (slot < 0) return HP_E_NOMEM;
    int r = alloc_box(b, slot, na, na, na, nb, precision, 0);
    if (r) return r;
    int axis;
    double x0;
    int flip;
    hp_sod_of_seed(seed, &axis, &x0, &flip);
    const hp::BatchBox& X = b->h_boxes[slot];
    const long tot = (long)na * na * na;
    long grid = std::min<long>((tot + 255) / 256, 16384);
    hipLaunchKernelGGL(sod_ic_kernel, dim3((unsigned)grid), dim3(256), 0, b->stream, X.f[0][0], X.f[0][1], X.f[0][2],
                       X.f[0][3], X.f[0][4], n, nb, axis, x0, flip);
    if (hipGetLastError() != hipSuccess) { free_box(b, slot); return HP_E_HIP; }
    r = start_box(b, slot, 5.0 / 3.0, cour, 1e30, nullptr, nsteps);
    if (r) { free_box(b, slot); push_nbox(b); return r; }
    *id = slot;
    return HP_OK;
}

int hp_sod_of_seed(uint64_t seed, int* axis, double* x0, int* flip)
{
    if (!axis || !x0 || !flip) return HP_E_ARG;
    const uint64_t h = splitmix64(seed);
    *axis = (int)(h % 3u);
    *x0 = 0.25 + 0.5 * (double)((h >> 8) & 0xFFFFu) / 65536.0;
    *flip = (int)((h >> 40) & 1u);
    return HP_OK;
}

int hp_batch_remove(hp_batch* b, int id)
{
    if (!b || id < 0 || id >= b->maxb || !b->h_boxes[id].active) return HP_E_ARG;
    HPCHK(hipStreamSynchronize(b->stream));
    b->h_boxes[id].active = 0;
    if (push_box(b, id)) return HP_E_HIP;
    HPCHK(hipStreamSynchronize(b->stream));      // the descriptor copy reads h_boxes[id]
    free_box(b, id);
    return push_nbox(b);
}

// hp_box_status of every slot from copies of the first n device statuses
// and scalars.
static int decode_status(hp_batch* b, const hp::DevStatus* st, const hp::Params* prm, int n, hp_box_status* out,
                         int* nactive)
{
    int na = 0;
    for (int i = 0; i < b->maxb; ++i) {
        hp_box_status s;
        memset(&s, 0, sizeof(s));
        if (i < n && b->h_boxes[i].active) {
            s.active = 1;
            s.done = st[i].done;
            s.steps_done = st[i].nstep_done;
            s.t = st[i].comoving ? st[i].cosmo.t : st[i].t;
            s.dt = prm[i].dt;
            s.newton_iters_max = st[i].newton_max;
            s.newton_nonconv = (int64_t)st[i].newton_nonconv;
            s.neg_state_clamps = (int64_t)st[i].clamps;
            ++na;
        }
        b->h_status[i] = s;
        if (out) out[i] = s;
    }
    if (nactive) *nactive = na;
    return HP_OK;
}

int hp_batch_poll(hp_batch* b, hp_box_status* out, int* nactive)
{
    if (!b) return HP_E_ARG;
    std::vector<hp::DevStatus> st(b->maxb);
    std::vector<hp::Params> prm(b->maxb);
    if (b->nbox > 0) {
        HPCHK(hipMemcpyAsync(st.data(), b->d_st, sizeof(hp::DevStatus) * b->nbox, hipMemcpyDeviceToHost, b->stream));
        HPCHK(hipMemcpyAsync(prm.data(), b->d_prm, sizeof(hp::Params) * b->nbox, hipMemcpyDeviceToHost, b->stream));
    }
    HPCHK(hipStreamSynchronize(b->stream));
    return decode_status(b, st.data(), prm.data(), b->nbox, out, nactive);
}



int hp_batch_download(hp_batch* b, int id, double* p, double* rho, double* vx, double* vy, double* vz)
{
    if (!b || id < 0 || id >= b->maxb || !b->h_boxes[id].active) return HP_E_ARG;
    if (!p || !rho || !vx || !vy || !vz) return HP_E_ARG;
    const hp::BatchBox& X = b->h_boxes[id];
    long long n = 0;
    HPCHK(hipMemcpyAsync(&n, &X.st->nstep_done, sizeof(n), hipMemcpyDeviceToHost, b->stream));
    HPCHK(hipStreamSynchronize(b->stream));
    const int cur = hp::batch_current(X.nsets, n);
    double* const f[5] = {X.f[cur][0], X.f[cur][1], X.f[cur][2], X.f[cur][3], X.f[cur][4]};
    HPCHK(hp::launch_fill(f, X.nx, X.ny, X.nz, X.nb, 0, b->stream));
    double* dst[5] = {p, rho, vx, vy, vz};
    const size_t bytes = (size_t)X.nx * X.ny * X.nz * sizeof(double);
    for (int k = 0; k < 5; ++k) HPCHK(hipMemcpyAsync(dst[k], f[k], bytes, hipMemcpyDeviceToHost, b->stream));
    HPCHK(hipStreamSynchronize(b->stream));
    return HP_OK;
}

int hp_batch_step(hp_batch* b, int nsteps)
{
    if (!b || nsteps < 0) return HP_E_ARG;
    const unsigned combos = combos_of(b);
    if (!combos) return HP_OK;
    const int grid = persistent_grid(b);
    if (b->use_graph) {
        if (!b->gexec || b->gcombos != combos || b->ggrid != grid) {
            int r = capture_step(b, combos, grid);
            if (r) return r;
        }
        for (int i = 0; i < nsteps; ++i) HPCHK(hipGraphLaunch(b->gexec, b->stream));
        return HP_OK;
    }
    for (int i = 0; i < nsteps; ++i) {
        int r = enqueue_step(b, combos, grid);
        if (r) return r;
    }
    return HP_OK;
}

int hp_batch_set_newton_limit(hp_batch* b, int id, int nmax)
{
    if (!b || id < 0 || id >= b->maxb || !b->h_boxes[id].active || nmax < 0) return HP_E_ARG;
    HPCHK(hipMemcpyAsync(&b->h_boxes[id].st->newton_cap, &nmax, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HPCHK(hipStreamSynchronize(b->stream));
    return HP_OK;
}

int hp_batch_sync(hp_batch* b)
{
    if (!b) return HP_E_ARG;
    HPCHK(hipStreamSynchronize(b->stream));
    return HP_OK;
}

int hp_batch_use_graph(hp_batch* b, int on)
{
    if (!b) return HP_E_ARG;
    HPCHK(hipStreamSynchronize(b->stream));
    b->use_graph = on != 0;
    if (!b->use_graph) drop_graph(b);
    return HP_OK;
}

int hp_batch_set_tuning(hp_batch* b, const hp_batch_tuning* t)
{
    if (!b || !t) return HP_E_ARG;
    if (t->x_chunks < 0 || t->x_chunks > 4096 || t->yz_segments < 0 || t->yz_segments > 64 ||
        t->waves_per_cu < 0 || t->waves_per_cu > 64)
        return HP_E_ARG;
    HPCHK(hipStreamSynchronize(b->stream));
    b->tune = *t;
    for (int i = 0; i < b->maxb; ++i) {
        if (!b->h_boxes[i].active) continue;
        set_units(b->h_boxes[i], b->tune);
        if (push_box(b, i)) return HP_E_HIP;
    }
    HPCHK(hipStreamSynchronize(b->stream));
    return HP_OK;
}

size_t hp_batch_bytes_used(const hp_batch* b) { return b ? b->used : 0; }

int hp_batch_get_tuning(hp_batch* b, hp_batch_tuning* t)
{
    if (!b || !t) return HP_E_ARG;
    *t = b->tune;
    return HP_OK;
}

// Online coordinate search over the tiling: every trial runs real steps of
// the batch (the tiling does not change any result, only the work split),
// timed with HIP events; each parameter in turn is set to its fastest value.
int hp_batch_autotune(hp_batch* b, int steps_per_trial, hp_batch_tuning* best, double* best_ms_per_step,
                      int* steps_run)
{
    if (!b || steps_per_trial < 1) return HP_E_ARG;
    static const int XC[] = {1, 2, 4, 8, 16};
    static const int YS[] = {1, 2, 4};
    static const int WP[] = {4, 8, 16};
    const bool graph_was = b->use_graph;      // trials run without graph capture, so that
    if (graph_was) hp_batch_use_graph(b, 0);  // host-side capture time stays out of the timings
    hp_batch_tuning cur = b->tune;
    if (cur.x_chunks == 0) cur.x_chunks = 4;
    if (cur.yz_segments == 0) cur.yz_segments = 1;
    if (cur.waves_per_cu == 0) cur.waves_per_cu = 8;
    hipEvent_t e0, e1;
    HPCHK(hipEventCreate(&e0));
    HPCHK(hipEventCreate(&e1));
    int run = 0;
    double best_ms = 1e300;
    auto trial = [&](const hp_batch_tuning& t, double& ms) -> int {
        int r = hp_batch_set_tuning(b, &t);
        if (r) return r;
        if (hipEventRecord(e0, b->stream) != hipSuccess) return HP_E_HIP;
        r = hp_batch_step(b, steps_per_trial);
        if (r) return r;
        if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return HP_E_HIP;
        float f = 0.f;
        if (hipEventElapsedTime(&f, e0, e1) != hipSuccess) return HP_E_HIP;
        ms = (double)f / steps_per_trial;
        run += steps_per_trial;
        return HP_OK;
    };
    // Each trial advances the boxes, and a step's cost drifts as the fields
    // evolve (more Newton iterations as structure forms), so the candidates of
    // a parameter are timed in order and again in reverse order and each is
    // scored by the mean of its two times: a drift linear in the step count
    // adds the same to every candidate's mean.
    int rc = HP_OK;
    for (int param = 0; param < 3 && rc == HP_OK; ++param) {
        const int* cand = param == 0 ? XC : (param == 1 ? YS : WP);
        const int nc = param == 0 ? 5 : 3;
        double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < 2 * nc && rc == HP_OK; ++k) {
            const int c = k < nc ? k : 2 * nc - 1 - k;
            hp_batch_tuning t = cur;
            if (param == 0) t.x_chunks = cand[c];
            else if (param == 1) t.yz_segments = cand[c];
            else t.waves_per_cu = cand[c];
            double ms = 0.0;
            rc = trial(t, ms);
            sum[c] += ms;
        }
        if (rc) break;
        int pick = 0;
        for (int c = 1; c < nc; ++c)
            if (sum[c] < sum[pick]) pick = c;
        if (param == 0) cur.x_chunks = cand[pick];
        else if (param == 1) cur.yz_segments = cand[pick];
        else cur.waves_per_cu = cand[pick];
        best_ms = 0.5 * sum[pick];
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (graph_was) hp_batch_use_graph(b, 1);
    if (rc) return rc;
    rc = hp_batch_set_tuning(b, &cur);
    if (best) *best = cur;
    if (best_ms_per_step) *best_ms_per_step = best_ms;
    if (steps_run) *steps_run = run;
    return rc;
}

}  // extern "C"

namespace hp {

int batch_status_async(hp_batch* b, int slot)
{
    if (!b || slot < 0 || slot > 1) return HP_E_ARG;
    if (!b->pin_st[slot]) {
        if (hipHostMalloc((void**)&b->pin_st[slot], sizeof(hp::DevStatus) * b->maxb, hipHostMallocDefault) !=
                hipSuccess ||
            hipHostMalloc((void**)&b->pin_prm[slot], sizeof(hp::Params) * b->maxb, hipHostMallocDefault) !=
                hipSuccess ||
            hipEventCreateWithFlags(&b->pin_ev[slot], hipEventDisableTiming) != hipSuccess)
            return HP_E_NOMEM;
    }
    const int n = b->nbox;
    b->pin_n[slot] = n;
    if (n > 0) {
        HPCHK(hipMemcpyAsync(b->pin_st[slot], b->d_st, sizeof(hp::DevStatus) * n, hipMemcpyDeviceToHost, b->stream));
        HPCHK(hipMemcpyAsync(b->pin_prm[slot], b->d_prm, sizeof(hp::Params) * n, hipMemcpyDeviceToHost, b->stream));
    }
    HPCHK(hipEventRecord(b->pin_ev[slot], b->stream));
    return HP_OK;
}

int batch_status_wait(hp_batch* b, int slot, hp_box_status* out, int* nactive)
{
    if (!b || slot < 0 || slot > 1 || !b->pin_ev[slot]) return HP_E_ARG;
    HPCHK(hipEventSynchronize(b->pin_ev[slot]));
    return decode_status(b, b->pin_st[slot], b->pin_prm[slot], b->pin_n[slot], out, nactive);
}

}  // namespace hp

