// svpf_kernels.hip -- gfx950 kernels of the SVPF step.
//
// svpf_batched_kernel: a persistent workgroup pool pulls filters from an
// atomic work queue (filters pre-sorted by decreasing N^2 cost) and advances
// each one through ticks [t0, t1) with its particles resident in registers and
// LDS.  A single filter is the B = 1 case.  Tile shape (rows per lane R,
// column split CS) is chosen per filter on the host from the tile table.
#include <hip/hip_runtime.h>

#include "svpf_engine.hpp"
#include "svpf_kernels.h"

using namespace svpf;

namespace {

template <int BS, int R, int CS>
__device__ __forceinline__ void run_filter(char* lds, int nmax, const svpf_fparams* params, svpf_fstate* fstate,
                           const svpf_fdesc* desc, int f, float* pool_h, float* pool_hp, float* pool_v, int t0,
                           int t1) {
    const svpf_fdesc* D = &desc[f];
    const svpf_fparams* P = &params[D->param_idx];
    FilterIO io{P, &fstate[f], D, pool_h + D->off, pool_hp + D->off, pool_v + D->off};
    Engine<BS, R, CS> e;
    e.bind(lds, P->n, nmax);
    e.load(io);
    for (int t = t0; t < t1; ++t) e.tick(io, t);
    e.store(io);
}

}  // namespace

template <int BS, int R, int CS>
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
            run_filter<BS, R, CS>(lds, nmax, params, fstate, desc, order[blockIdx.x], pool_h, pool_hp, pool_v, t0, t1);
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(queue, 1);
        __syncthreads();
        const int k = next;
        __syncthreads();
        if (k >= nf) break;  // every workgroup reaches this once the queue is drained
        run_filter<BS, R, CS>(lds, nmax, params, fstate, desc, order[k], pool_h, pool_hp, pool_v, t0, t1);
    }
}

__global__ void svpf_queue_reset_kernel(int* queue) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *queue = 0;
}

#define SVPF_SET_LDS_ATTR(BSV, RV, CSV)                                                                      \
    do {                                                                                                     \
        hipError_t e_ = hipFuncSetAttribute((const void*)svpf_batched_kernel<BSV, RV, CSV>,                 \
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);               \
        if (e_ != hipSuccess) return e_;                                                                     \
    } while (0)

// Lets every engine instance use up to 159 KiB of dynamic LDS (gfx950 has 160 KiB per CU).
extern "C" hipError_t svpf_kernels_init(void) {
    static hipError_t done = hipErrorNotReady;
    if (done == hipSuccess) return done;
    const int bytes = 160 * 1024 - 1024;
    SVPF_SET_LDS_ATTR(256, 1, 4);
    SVPF_SET_LDS_ATTR(256, 1, 2);
    SVPF_SET_LDS_ATTR(256, 1, 1);
    SVPF_SET_LDS_ATTR(256, 2, 1);
    SVPF_SET_LDS_ATTR(256, 4, 1);
    SVPF_SET_LDS_ATTR(256, 8, 1);
    SVPF_SET_LDS_ATTR(256, 2, 2);
    SVPF_SET_LDS_ATTR(256, 4, 2);
    SVPF_SET_LDS_ATTR(512, 1, 4);
    SVPF_SET_LDS_ATTR(512, 1, 2);
    SVPF_SET_LDS_ATTR(512, 1, 1);
    SVPF_SET_LDS_ATTR(512, 2, 1);
    SVPF_SET_LDS_ATTR(512, 4, 1);
    SVPF_SET_LDS_ATTR(512, 8, 1);
    SVPF_SET_LDS_ATTR(512, 2, 2);
    SVPF_SET_LDS_ATTR(512, 4, 2);
    SVPF_SET_LDS_ATTR(1024, 1, 1);
    SVPF_SET_LDS_ATTR(1024, 2, 1);
    SVPF_SET_LDS_ATTR(1024, 4, 1);
    SVPF_SET_LDS_ATTR(1024, 1, 2);
    SVPF_SET_LDS_ATTR(1024, 1, 4);
    done = hipSuccess;
    return done;
}

__global__ void svpf_init_particles_kernel(float* h, float* hp, float* v, int n, float mu, float sd, uint32_t k0,
                                           uint32_t k1) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o[4];
    philox4x32_10((uint32_t)i, 0u, SVPF_PURPOSE_INIT, 0u, k0, k1, o);
    float z = normal_from_bits(o[0], o[1]);
    float x = clamp_logvol(mu + sd * z);
    h[i] = x;
    hp[i] = x;
    v[i] = 0.0f;
}

extern "C" hipError_t svpf_launch_init(float* h, float* hp, float* v, int n, float mu, float sd, uint64_t seed,
                                       hipStream_t stream) {
    if (n <= 0) return hipErrorInvalidValue;
    svpf_init_particles_kernel<<<(n + 255) / 256, 256, 0, stream>>>(h, hp, v, n, mu, sd, (uint32_t)seed,
                                                                     (uint32_t)(seed >> 32));
    return hipGetLastError();
}

extern "C" size_t svpf_lds_bytes(int nmax) { return LdsLayout::bytes(nmax); }

#define SVPF_OCC_TILE(BSV, RV, CSV) fn = (const void*)svpf_batched_kernel<BSV, RV, CSV>

// Workgroups of the given tile resident per CU for filters of up to nmax particles (0: unknown tile).
extern "C" int svpf_batched_wg_per_cu(int bs, int tile, int nmax) {
    if (svpf_kernels_init() != hipSuccess) return 0;
    const void* fn = nullptr;
    if (bs == 256) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_OCC_TILE(256, 1, 4); break;
            case SVPF_TILE_R1_C2: SVPF_OCC_TILE(256, 1, 2); break;
            case SVPF_TILE_R1_C1: SVPF_OCC_TILE(256, 1, 1); break;
            case SVPF_TILE_R2_C1: SVPF_OCC_TILE(256, 2, 1); break;
            case SVPF_TILE_R4_C1: SVPF_OCC_TILE(256, 4, 1); break;
            case SVPF_TILE_R8_C1: SVPF_OCC_TILE(256, 8, 1); break;
            case SVPF_TILE_R2_C2: SVPF_OCC_TILE(256, 2, 2); break;
            case SVPF_TILE_R4_C2: SVPF_OCC_TILE(256, 4, 2); break;
            default: return 0;
        }
    } else if (bs == 512) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_OCC_TILE(512, 1, 4); break;
            case SVPF_TILE_R1_C2: SVPF_OCC_TILE(512, 1, 2); break;
            case SVPF_TILE_R1_C1: SVPF_OCC_TILE(512, 1, 1); break;
            case SVPF_TILE_R2_C1: SVPF_OCC_TILE(512, 2, 1); break;
            case SVPF_TILE_R4_C1: SVPF_OCC_TILE(512, 4, 1); break;
            case SVPF_TILE_R8_C1: SVPF_OCC_TILE(512, 8, 1); break;
            case SVPF_TILE_R2_C2: SVPF_OCC_TILE(512, 2, 2); break;
            case SVPF_TILE_R4_C2: SVPF_OCC_TILE(512, 4, 2); break;
            default: return 0;
        }
    } else if (bs == 1024) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_OCC_TILE(1024, 1, 4); break;
            case SVPF_TILE_R1_C2: SVPF_OCC_TILE(1024, 1, 2); break;
            case SVPF_TILE_R1_C1: SVPF_OCC_TILE(1024, 1, 1); break;
            case SVPF_TILE_R2_C1: SVPF_OCC_TILE(1024, 2, 1); break;
            case SVPF_TILE_R4_C1: SVPF_OCC_TILE(1024, 4, 1); break;
            default: return 0;
        }
    } else {
        return 0;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, bs, LdsLayout::bytes(nmax)) != hipSuccess) return 0;
    return nb;
}

extern "C" int svpf_tile_capacity(int tile, int bs) {
    if (bs != 256 && bs != 512 && bs != 1024) return 0;
    if (bs == 1024 && tile > SVPF_TILE_R4_C1) return 0;  // not instantiated at 1024 threads
    switch (tile) {
        case SVPF_TILE_R1_C4: return bs / 4;
        case SVPF_TILE_R1_C2: return bs / 2;
        case SVPF_TILE_R1_C1: return bs;
        case SVPF_TILE_R2_C1: return 2 * bs;
        case SVPF_TILE_R4_C1: return 4 * bs;
        case SVPF_TILE_R8_C1: return 8 * bs;
        case SVPF_TILE_R2_C2: return bs;
        case SVPF_TILE_R4_C2: return 2 * bs;
        default: return 0;
    }
}

#define SVPF_LAUNCH_TILE(BSV, RV, CSV)                                                                   \
    svpf_batched_kernel<BSV, RV, CSV><<<grid, BSV, lds, stream>>>(params, fstate, desc, pool_h, pool_hp, pool_v, \
                                                                 order, nf, queue, nmax, t0, t1)

extern "C" hipError_t svpf_launch_batched(int bs, int tile, int grid, const svpf_fparams* params,
                                          svpf_fstate* fstate, const svpf_fdesc* desc, float* pool_h,
                                          float* pool_hp, float* pool_v, const int* order, int nf, int* queue,
                                          int nmax, int t0, int t1, hipStream_t stream) {
    size_t lds = LdsLayout::bytes(nmax);
    if (queue) svpf_queue_reset_kernel<<<1, 64, 0, stream>>>(queue);
    if (bs == 256) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_LAUNCH_TILE(256, 1, 4); break;
            case SVPF_TILE_R1_C2: SVPF_LAUNCH_TILE(256, 1, 2); break;
            case SVPF_TILE_R1_C1: SVPF_LAUNCH_TILE(256, 1, 1); break;
            case SVPF_TILE_R2_C1: SVPF_LAUNCH_TILE(256, 2, 1); break;
            case SVPF_TILE_R4_C1: SVPF_LAUNCH_TILE(256, 4, 1); break;
            case SVPF_TILE_R8_C1: SVPF_LAUNCH_TILE(256, 8, 1); break;
            case SVPF_TILE_R2_C2: SVPF_LAUNCH_TILE(256, 2, 2); break;
            case SVPF_TILE_R4_C2: SVPF_LAUNCH_TILE(256, 4, 2); break;
            default: return hipErrorInvalidValue;
        }
    } else if (bs == 512) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_LAUNCH_TILE(512, 1, 4); break;
            case SVPF_TILE_R1_C2: SVPF_LAUNCH_TILE(512, 1, 2); break;
            case SVPF_TILE_R1_C1: SVPF_LAUNCH_TILE(512, 1, 1); break;
            case SVPF_TILE_R2_C1: SVPF_LAUNCH_TILE(512, 2, 1); break;
            case SVPF_TILE_R4_C1: SVPF_LAUNCH_TILE(512, 4, 1); break;
            case SVPF_TILE_R8_C1: SVPF_LAUNCH_TILE(512, 8, 1); break;
            case SVPF_TILE_R2_C2: SVPF_LAUNCH_TILE(512, 2, 2); break;
            case SVPF_TILE_R4_C2: SVPF_LAUNCH_TILE(512, 4, 2); break;
            default: return hipErrorInvalidValue;
        }
    } else if (bs == 1024) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_LAUNCH_TILE(1024, 1, 4); break;
            case SVPF_TILE_R1_C2: SVPF_LAUNCH_TILE(1024, 1, 2); break;
            case SVPF_TILE_R1_C1: SVPF_LAUNCH_TILE(1024, 1, 1); break;
            case SVPF_TILE_R2_C1: SVPF_LAUNCH_TILE(1024, 2, 1); break;
            case SVPF_TILE_R4_C1: SVPF_LAUNCH_TILE(1024, 4, 1); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cooperative latency kernel: one filter over G workgroups (see svpf_coop.hpp).
// Workgroups whose index is not a multiple of `stride` exit at once, so with
// stride 8 the G participants share one XCD (workgroups are dealt to the XCDs
// round-robin) and exchange through one L2.
// ---------------------------------------------------------------------------
#include "svpf_coop.hpp"

template <int BS, int CS>
__global__ __launch_bounds__(BS) void svpf_coop_kernel(const svpf_fparams* __restrict__ params,
                                                       svpf_fstate* __restrict__ fstate,
                                                       const svpf_fdesc* __restrict__ desc, float* pool_h,
                                                       float* pool_hp, float* pool_v, uint64_t* xbuf, int G,
                                                       int stride, int t0, int t1) {
    extern __shared__ __align__(16) char lds[];
    if ((int)blockIdx.x % stride) return;
    const int g = (int)blockIdx.x / stride;
    if (g >= G) return;
    const svpf_fdesc* D = desc;
    const svpf_fparams* P = &params[D->param_idx];
    CoopIO io{P, fstate, D, pool_h + D->off, pool_hp + D->off, pool_v + D->off, xbuf, G, g};
    CoopEngine<BS, CS> e;
    e.bind(lds, P->n, G, g);
    if (!e.run(io, t0, t1) && threadIdx.x == 0) fstate->status = 2;
}

#define SVPF_COOP_CASE(CSV)                                                                                  \
    case CSV:                                                                                                \
        fn = (const void*)svpf_coop_kernel<256, CSV>;                                                        \
        lds = CoopEngine<256, CSV>::lds_bytes(n);                                                            \
        break;

extern "C" int svpf_coop_rows(int cs) { return 256 / cs; }

#ifdef SVPF_TRACE
__device__ unsigned long long svpf_trace_buf[256 * 64];
extern "C" hipError_t svpf_trace_dump(unsigned long long* out) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(svpf_trace_buf), sizeof(unsigned long long) * 256 * 64);
}
#endif

extern "C" size_t svpf_coop_lds_bytes(int n) { return CoopEngine<256, 1>::lds_bytes(n); }

extern "C" hipError_t svpf_launch_coop(int cs, int G, int stride, const svpf_fparams* params, svpf_fstate* fstate,
                                       const svpf_fdesc* desc, float* pool_h, float* pool_hp, float* pool_v,
                                       uint64_t* xbuf, int n, int t0, int t1, int cooperative, hipStream_t stream) {
    const void* fn = nullptr;
    size_t lds = 0;
    switch (cs) {
        SVPF_COOP_CASE(1)
        SVPF_COOP_CASE(2)
        SVPF_COOP_CASE(4)
        SVPF_COOP_CASE(8)
        SVPF_COOP_CASE(16)
        SVPF_COOP_CASE(32)
        default: return hipErrorInvalidValue;
    }
    if (G < 1 || stride < 1 || (long)G * (256 / cs) < n || G * 8 > n) return hipErrorInvalidValue;
    if (lds > 159 * 1024) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    void* args[] = {(void*)&params, (void*)&fstate, (void*)&desc, (void*)&pool_h, (void*)&pool_hp, (void*)&pool_v,
                    (void*)&xbuf, (void*)&G, (void*)&stride, (void*)&t0, (void*)&t1};
    dim3 grid(G * stride), block(256);
    if (cooperative) return hipLaunchCooperativeKernel(fn, grid, block, args, (unsigned)lds, stream);
    return hipLaunchKernel(fn, grid, block, args, lds, stream);
}
