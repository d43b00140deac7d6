// hp_batch_kernels.hip -- persistent sweep kernels over a batch of boxes
// (throughput arithmetic of hp_fast.h, FP contraction on).  Each wave takes
// work units u = wave, wave + nwaves, ... of one plan table, finds the box by
// binary search over the table's unit prefix sums, and runs the same per-box
// sweep bodies as the single-box kernels (x_run / stream_run).  Every box
// uses the comoving form of the tracing (at = 1, dat = 0 for a static box);
// gravity is not supported in a batch.
#include "hp_batch.h"
#include "hp_sweep.h"

namespace hp {

__device__ __forceinline__ int table_find(const BatchTable& T, int u)
{
    int lo = 0, hi = T.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T.start[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// v with every 32-bit word taken from the wave's first active lane.  The
// per-box descriptors and scalars are wave-uniform, but the kernel reads them
// after its own counter atomics, through pointers it also writes, so the
// compiler loads them per lane; this returns them to scalar registers.
template <class T> __device__ __forceinline__ T wave_uniform(const T& v)
{
    static_assert(sizeof(T) % 4 == 0, "whole words");
    int w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

__device__ __forceinline__ void batch_view(const BatchBox& X, int slot, BoxView& V, Epilogue& E)
{
    int in, out, old;
    batch_sets(X.nsets, X.st->nstep_done, slot, in, out, old);
    in = __builtin_amdgcn_readfirstlane(in);
    out = __builtin_amdgcn_readfirstlane(out);
    old = __builtin_amdgcn_readfirstlane(old);
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        V.in[f] = X.f[in][f];
        V.out[f] = X.f[out][f];
    }
    V.gforce = nullptr;
    V.prm = X.prm;
    V.st = X.st;
    V.nx = X.nx; V.ny = X.ny; V.nz = X.nz; V.nb = X.nb;
    const bool exp = slot == 2 && old >= 0;
    E.old[0] = exp ? X.f[old][0] : nullptr;
    E.old[1] = exp ? X.f[old][2] : nullptr;
    E.old[2] = exp ? X.f[old][3] : nullptr;
    E.old[3] = exp ? X.f[old][4] : nullptr;
    E.fac = X.st->fac;
    E.afact = E.afact1 = E.afact2 = 0.0;
    if (exp) {
        load_factors(E);
        E.afact = wave_uniform(E.afact);
        E.afact1 = wave_uniform(E.afact1);
        E.afact2 = wave_uniform(E.afact2);
    }
    E.ghost_mask = 7;
    E.do_speed = slot == 2;
}

template <class M, class R, bool EXP>
__global__ void __launch_bounds__(256) batch_x(const BatchBox* __restrict__ boxes, const BatchPlan* __restrict__ plan,
                                               int key, int slot)
{
    const BatchTable& T = plan->t[key];
    const int total = T.total;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    // unit indices are wave-uniform; readfirstlane lets the table search and
    // the descriptor loads run on the scalar unit
    const int u0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    int cur = -1;                                    // box whose tallies t holds
    Tally t = tally0();
    for (int u = u0; u < total; u += nwaves) {
        const int i = __builtin_amdgcn_readfirstlane(table_find(T, u));
        const int b = __builtin_amdgcn_readfirstlane(T.box[i]);
        if (b != cur) {                              // one flush per box a wave visits
            if (cur >= 0) flush_counters(boxes[cur].st, t, slot == 2);
            t = tally0();
            cur = b;
        }
        const BatchBox& X = boxes[b];
        const int local = u - T.start[i];
        BoxView V;
        Epilogue E;
        batch_view(X, slot, V, E);
        const unsigned nchunks = x_chunks(X.nx, X.ny, X.nz, X.nb);
        const unsigned c0 = (unsigned)local * (unsigned)X.xcpu;
        const unsigned c1 = min(c0 + (unsigned)X.xcpu, nchunks);
        const AllScalars<R> A = wave_uniform(load_scalars<M, R>(X.prm, X.st));
        x_run<M, R, false, true, EXP>(V, A, E, c0, c1, t);
    }
    if (cur >= 0) flush_counters(boxes[cur].st, t, slot == 2);
}

// At least two waves per SIMD for the gamma = 5/3 forms and the mixed forms:
// left to itself the allocator gives the 5/3 FP64 z-sweep body 257 registers
// and the general-gamma mixed bodies 259, just above the 256 that two waves
// allow, and those kernels run at one wave per SIMD; held at two they fit
// with no spill.  (The general-gamma FP64 bodies need ~290; forced to two
// waves they spill about 1 KB per lane, so they are left to the allocator.)
template <class M, class R, int DIR, bool EXP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((M::G53 || sizeof(R) == 4) ? 2 : 1)))
batch_stream(const BatchBox* __restrict__ boxes, const BatchPlan* __restrict__ plan, int key, int slot)
{
    const BatchTable& T = plan->t[key];
    const int total = T.total;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    // unit indices are wave-uniform; readfirstlane lets the table search and
    // the descriptor loads run on the scalar unit
    const int u0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    int cur = -1;
    Tally t = tally0();
    for (int u = u0; u < total; u += nwaves) {
        const int i = __builtin_amdgcn_readfirstlane(table_find(T, u));
        const int b = __builtin_amdgcn_readfirstlane(T.box[i]);
        if (b != cur) {
            if (cur >= 0) flush_counters(boxes[cur].st, t, slot == 2);
            t = tally0();
            cur = b;
        }
        const BatchBox& X = boxes[b];
        const int local = u - T.start[i];
        BoxView V;
        Epilogue E;
        batch_view(X, slot, V, E);
        const int Nbb = (DIR == 2) ? X.nz - 2 * X.nb : X.ny - 2 * X.nb;
        const long npen = (long)(X.nx - 2 * X.nb) * Nbb;
        const int nseg = X.nseg[DIR - 2];
        const unsigned item = (unsigned)local * 64u + (unsigned)lane;   // items per box < 2^31
        const bool active = (long)item < npen * nseg;
        const unsigned it = active ? item : 0u;
        const int seg = (int)(it / (unsigned)npen);
        const long pen = (long)(it - (unsigned)seg * (unsigned)npen);
        const AllScalars<R> A = wave_uniform(load_scalars<M, R>(X.prm, X.st));
        stream_run<M, R, DIR, false, true, EXP>(V, A, E, pen, seg, X.seglen[DIR - 2], active, t);
    }
    if (cur >= 0) flush_counters(boxes[cur].st, t, slot == 2);
}

template <class M, class R, bool EXP>
static void launch_dir(const BatchBox* boxes, const BatchPlan* plan, int key, int slot, int dir, int grid,
                       hipStream_t s)
{
    if (dir == 1) hipLaunchKernelGGL((batch_x<M, R, EXP>), dim3(grid), dim3(256), 0, s, boxes, plan, key, slot);
    else if (dir == 2)
        hipLaunchKernelGGL((batch_stream<M, R, 2, EXP>), dim3(grid), dim3(256), 0, s, boxes, plan, key, slot);
    else hipLaunchKernelGGL((batch_stream<M, R, 3, EXP>), dim3(grid), dim3(256), 0, s, boxes, plan, key, slot);
}

template <class M>
static void launch_m(const BatchBox* boxes, const BatchPlan* plan, int key, int slot, int dir, int prec, int exp,
                     int grid, hipStream_t s)
{
    if (prec == PREC_MIXED) {
        if (exp) launch_dir<M, float, true>(boxes, plan, key, slot, dir, grid, s);
        else launch_dir<M, float, false>(boxes, plan, key, slot, dir, grid, s);
    } else {
        if (exp) launch_dir<M, double, true>(boxes, plan, key, slot, dir, grid, s);
        else launch_dir<M, double, false>(boxes, plan, key, slot, dir, grid, s);
    }
}

hipError_t launch_batch_sweep(const BatchBox* boxes, const BatchPlan* plan, int slot, int dir, int prec, int exp,
                              int g53, int grid, hipStream_t s)
{
    if (slot < 0 || slot > 2 || dir < 1 || dir > 3 || grid < 1) return hipErrorInvalidValue;
    const int key = table_key(slot, dir, prec, exp, g53);
    if (g53) launch_m<FastMath53>(boxes, plan, key, slot, dir, prec, exp, grid, s);
    else launch_m<FastMath>(boxes, plan, key, slot, dir, prec, exp, grid, s);
    return hipGetLastError();
}

}  // namespace hp
