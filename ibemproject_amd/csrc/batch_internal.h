// Internal interface between the C-ABI contexts (api.cpp) and the batched solver (batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

struct ibem3d_ctx;

namespace ibem_batch {

hipStream_t stream(ibem3d_ctx* ctx);
long size(ibem3d_ctx* ctx);
// Allocate every workspace a captured mixed solve of `mode` uses (eager: may allocate and sync).
int prepare(ibem3d_ctx* ctx, int mode);
// Enqueue on the context's streams, with no host synchronisation, allocation or memset (graph
// capturable): the assembly with load 0's displacement values, the mixed LU and its solve
// preparation, then for each load t < nloads (displacement values d_uvals + t * 3 NN): b (for
// t > 0 the right-hand-side-only assembly), K classic refinement steps from x = 0, x copied to
// d_xs + t * n and ||b - A x|| / ||b|| to d_rel[t]. d_part: 2 kDotBlocks doubles.
int enqueue(ibem3d_ctx* ctx, int mode, int K, int nloads, const double* d_uvals, double* d_xs,
            double* d_rel, double* d_part);
// After the enqueued work ran: the context holds the factors of `mode`, the system of the last load.
void mark_solved(ibem3d_ctx* ctx, int mode);
// Volume averages of a host solution x of the load whose prescribed displacements are uvals (3 NN;
// they become the context's known values): boundary stress and strain, eigenstrain average, volume.
int averages(ibem3d_ctx* ctx, const double* x, const double* uvals, double* sigma, double* eps,
             double* eps_star, double* V);
size_t dot_blocks();

}  // namespace ibem_batch
