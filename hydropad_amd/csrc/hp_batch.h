// hp_batch.h -- device-side layout of a batch of independent boxes advanced
// together: one descriptor per box, and per step a work plan that lists, for
// each (sweep slot, direction, precision, fused expansion), the boxes that
// need that sweep and the prefix sums of their work units.  The sweep
// kernels are persistent: a fixed grid walks the units of one plan table, so
// the launch sequence of a step does not depend on which boxes are present
// and can be captured once in a hipGraph.
#pragma once
#include "hp_kernels.h"

namespace hp {

constexpr int BATCH_MAXB = 1024;          // boxes per batch
constexpr int BATCH_NTABLE = 72;          // 3 slots x 3 directions x 2 precisions x expansion x gamma 5/3

__host__ __device__ constexpr int table_key(int slot, int dir, int prec, int exp, int g53)
{
    return (((slot * 3 + (dir - 1)) * 2 + prec) * 2 + exp) * 2 + g53;
}

struct BatchBox {
    double* f[3][5];      // field sets (the third only for comoving boxes)
    Params* prm;
    DevStatus* st;
    int nx, ny, nz, nb;
    int nsets;            // 2, or 3 when the step keeps its start state for the expansion
    int prec;             // PREC_FP64 or PREC_MIXED
    int g53;              // gamma = 5/3 scalars: the constant-folded kernels
    int active;           // slot in use
    int xcpu;             // x sweep: chunks of 58 cells per work unit
    int seglen[2];        // y, z sweeps: streaming segment length
    int nseg[2];          // y, z sweeps: segments per pencil
    int units[3];         // work units per direction
};

struct BatchTable {
    int n;                // boxes in the table
    int total;            // work units
    int box[BATCH_MAXB];
    int start[BATCH_MAXB];
};

struct BatchPlan {
    int nbox;             // descriptors in use (highest slot + 1)
    int pad[3];
    BatchTable t[BATCH_NTABLE];
};

// Field sets of a box at sweep `slot` (0..2) of its step number n+1 (n steps
// done).  Two sets ping-pong; with three, the step's start set S = n mod 3 is
// kept for the expansion: S -> S+1 -> S+2 -> S+1.
__host__ __device__ inline void batch_sets(int nsets, long long n, int slot, int& in, int& out, int& old)
{
    if (nsets == 2) {
        const long long k = 3 * n + slot;
        in = (int)(k & 1);
        out = (int)((k + 1) & 1);
        old = -1;
    } else {
        const int S = (int)(n % 3);
        const int ins[3] = {S, (S + 1) % 3, (S + 2) % 3};
        const int outs[3] = {(S + 1) % 3, (S + 2) % 3, (S + 1) % 3};
        in = ins[slot];
        out = outs[slot];
        old = S;
    }
}

// The set holding the current state of a box after n steps.
__host__ __device__ inline int batch_current(int nsets, long long n)
{
    return nsets == 2 ? (int)(n & 1) : (int)(n % 3);
}

hipError_t launch_batch_sweep(const BatchBox* boxes, const BatchPlan* plan, int slot, int dir, int prec, int exp,
                              int g53, int grid, hipStream_t s);

}  // namespace hp

struct hp_batch;
struct hp_box_status;

namespace hp {

// Status reads that do not stall the stream (the serving loop's form of
// hp_batch_poll): batch_status_async enqueues a copy of every box's status
// and scalars into pinned host buffer `slot` (0 or 1) behind the work queued
// so far and records an event; batch_status_wait waits for that event only
// and decodes the copy, while later queued steps keep the GPU busy.
int batch_status_async(hp_batch* b, int slot);
int batch_status_wait(hp_batch* b, int slot, hp_box_status* out, int* nactive);

}  // namespace hp
