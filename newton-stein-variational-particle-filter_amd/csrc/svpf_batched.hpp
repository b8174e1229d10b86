// svpf_batched.hpp -- the persistent batched kernel over tile shapes.
//
// A persistent workgroup pool pulls filters from an atomic work queue (filters
// pre-sorted by decreasing N^2 cost) and advances each one through ticks
// [t0, t1) with its particles resident in registers and LDS.  A single filter is
// the B = 1 case.  EXT selects the instantiations with the Heun step and the
// Student-t state; they live in their own translation unit (explicit
// instantiation of batched_kernel<true>) so the common kernels keep their
// register budget and the two sets compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "svpf_engine.hpp"
#include "svpf_kernels.h"

namespace svpf {

template <int BS, int R, int CS, bool EXT>
__device__ __forceinline__ void run_filter(char* lds, int nmax, const svpf_fparams* params, svpf_fstate* fstate,
                                           const svpf_fdesc* desc, int f, float* pool_h, float* pool_hp,
                                           float* pool_v, int t0, int t1) {
    const svpf_fdesc* D = &desc[f];
    const svpf_fparams* P = &params[D->param_idx];
    // a filter with the Heun step or the Student-t state in a kernel without them:
    // flagged (status 1) for the host to report, its ticks run without them
    if (!EXT && threadIdx.x == 0 && (P->flags & (SVPF_F_HEUN | SVPF_F_STUDENT_STATE))) fstate[f].status = 1;
    FilterIO io{P, &fstate[f], D, pool_h + D->off, pool_hp + D->off, pool_v + D->off};
    Engine<BS, R, CS, EXT> e;
    e.bind(lds, P->n, nmax);
    e.load(io);
    for (int t = t0; t < t1; ++t) e.tick(io, t);
    e.store(io);
}

template <int BS, int R, int CS, bool EXT>
__global__ __launch_bounds__(BS) void svpf_batched_kernel(const svpf_fparams* __restrict__ params,
                                                          svpf_fstate* __restrict__ fstate,
                                                          const svpf_fdesc* __restrict__ desc,
                                                          float* __restrict__ pool_h, float* __restrict__ pool_hp,
                                                          float* __restrict__ pool_v, const int* __restrict__ order,
                                                          int nf, int* __restrict__ queue, int nmax, int t0, int t1) {
    extern __shared__ __align__(16) char lds[];
    __shared__ int next;
    if (queue == nullptr) {  // static assignment: workgroup b runs order[b]
        if ((int)blockIdx.x < nf)
            run_filter<BS, R, CS, EXT>(lds, nmax, params, fstate, desc, order[blockIdx.x], pool_h, pool_hp, pool_v,
                                       t0, t1);
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(queue, 1);
        __syncthreads();
        const int k = next;
        __syncthreads();
        if (k >= nf) break;  // every workgroup reaches this once the queue is drained
        run_filter<BS, R, CS, EXT>(lds, nmax, params, fstate, desc, order[k], pool_h, pool_hp, pool_v, t0, t1);
    }
}

// The kernel of a (block size, tile) pair, nullptr for an unknown one.  Block
// sizes 256 and 512 take every tile, 1024 the tiles up to R4_C1.
template <bool EXT>
const void* batched_kernel(int bs, int tile) {
#define SVPF_K(BSV, RV, CSV) return (const void*)svpf_batched_kernel<BSV, RV, CSV, EXT>
#define SVPF_TILES(BSV)                                 \
    switch (tile) {                                     \
        case SVPF_TILE_R1_C4: SVPF_K(BSV, 1, 4);        \
        case SVPF_TILE_R1_C2: SVPF_K(BSV, 1, 2);        \
        case SVPF_TILE_R1_C1: SVPF_K(BSV, 1, 1);        \
        case SVPF_TILE_R2_C1: SVPF_K(BSV, 2, 1);        \
        case SVPF_TILE_R4_C1: SVPF_K(BSV, 4, 1);        \
        case SVPF_TILE_R8_C1: SVPF_K(BSV, 8, 1);        \
        case SVPF_TILE_R2_C2: SVPF_K(BSV, 2, 2);        \
        case SVPF_TILE_R4_C2: SVPF_K(BSV, 4, 2);        \
        default: return nullptr;                        \
    }
    if (bs == 256) {
        SVPF_TILES(256)
    } else if (bs == 512) {
        SVPF_TILES(512)
    } else if (bs == 1024) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_K(1024, 1, 4);
            case SVPF_TILE_R1_C2: SVPF_K(1024, 1, 2);
            case SVPF_TILE_R1_C1: SVPF_K(1024, 1, 1);
            case SVPF_TILE_R2_C1: SVPF_K(1024, 2, 1);
            case SVPF_TILE_R4_C1: SVPF_K(1024, 4, 1);
            default: return nullptr;
        }
    }
    return nullptr;
#undef SVPF_TILES
#undef SVPF_K
}

extern template const void* batched_kernel<false>(int, int);
extern template const void* batched_kernel<true>(int, int);

}  // namespace svpf
