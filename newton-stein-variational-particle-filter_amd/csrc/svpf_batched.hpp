// svpf_batched.hpp -- the persistent batched kernel over tile shapes.
//
// A persistent workgroup pool pulls work items from an atomic queue: item k is
// tick t0 + k / nf of filter order[k % nf] (filters pre-sorted by decreasing N^2
// cost), so the ticks of all filters interleave and the pool stays busy to the
// end of the launch -- whole filters as items left up to a filter's T ticks of
// idle tail.  An item loads the filter's particles into registers and LDS, runs
// one tick and stores them; it first waits for the item of the filter's
// previous tick (k - nf), which a running workgroup took earlier, through the
// filter's progress word.  A single filter is the B = 1 case (no queue: the
// workgroup advances it through all ticks).  EXT selects the instantiations with the Heun step and the
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
                                                          int nf, int* __restrict__ queue, int* __restrict__ progress,
                                                          int nmax, int t0, int t1) {
    extern __shared__ __align__(16) char lds[];
    __shared__ int next;
    if (queue == nullptr) {  // static assignment: workgroup b runs order[b]
        if ((int)blockIdx.x < nf)
            run_filter<BS, R, CS, EXT>(lds, nmax, params, fstate, desc, order[blockIdx.x], pool_h, pool_hp, pool_v,
                                       t0, t1);
        return;
    }
    constexpr long kSpinLimit = 1L << 24;  // polls of a wait for the previous tick (about 10 s)
    const int items = nf * (t1 - t0);
    // The item finished last: thread 0 publishes its filter's progress together with the
    // next pop, after the barrier that ends the item (every thread's stores fenced).  A
    // divergent store at the end of the loop body instead was merged with the pop across
    // the back edge, which let lane 0 leave the loop while the other lanes repeated the item.
    int done_slot = -1, done_ticks = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            if (done_slot >= 0)
                __hip_atomic_store(progress + done_slot, done_ticks, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            next = atomicAdd(queue, 1);
        }
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (k >= items) break;  // every workgroup reaches this once the queue is drained
        const int slot = k % nf, tt = k / nf;
        if (tt > 0) {
            // progress[slot] counts the filter's finished ticks of this launch; the item of
            // its previous tick was taken before this one, by a workgroup that is running
            if (threadIdx.x == 0) {
                long spins = 0;
                while (__hip_atomic_load(progress + slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < tt) {
                    if (++spins > kSpinLimit) {
                        fstate[order[slot]].status |= 2;  // reported by the host; the tick runs regardless
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __syncthreads();
        }
        run_filter<BS, R, CS, EXT>(lds, nmax, params, fstate, desc, order[slot], pool_h, pool_hp, pool_v, t0 + tt,
                                   t0 + tt + 1);
        __threadfence();  // this workgroup's stores of the filter's state reach L2 before the release
        __syncthreads();
        done_slot = slot;
        done_ticks = tt + 1;
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
