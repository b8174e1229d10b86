// svpf_kernels.hip -- gfx950 kernels of the SVPF step.
//
// svpf_batched_kernel (svpf_batched.hpp): a persistent workgroup pool advancing
// filters tick by tick, one workgroup per filter.  The tile shape (rows per lane
// R, column split CS) is chosen per filter on the host from the tile table; a
// tile id with SVPF_TILE_EXT set selects the instantiation with the Heun step
// and the Student-t state (built in svpf_kernels_ext.hip).
#include <hip/hip_runtime.h>

#include "svpf_batched.hpp"

using namespace svpf;

namespace svpf {
template const void* batched_kernel<false>(int, int);
}

// The work queue's counter and the nf progress words of a batched launch start at 0.
__global__ void svpf_queue_reset_kernel(int* queue, int* progress, int nf) {
    if (threadIdx.x == 0) *queue = 0;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) progress[i] = 0;
}

// Lets every engine instance use up to 159 KiB of dynamic LDS (gfx950 has 160 KiB per CU).
extern "C" hipError_t svpf_kernels_init(void) {
    static hipError_t done = hipErrorNotReady;
    if (done == hipSuccess) return done;
    const int bytes = 160 * 1024 - 1024;
    for (int ext = 0; ext < 2; ++ext)
        for (int bs = 256; bs <= 1024; bs *= 2)
            for (int tile = 0; tile < SVPF_NUM_TILES; ++tile) {
                const void* fn = ext ? batched_kernel<true>(bs, tile) : batched_kernel<false>(bs, tile);
                if (!fn) continue;
                hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
                if (e != hipSuccess) return e;
            }
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

namespace {
const void* batched_fn(int bs, int tile) {
    if (tile & ~(SVPF_TILE_EXT | 0xff)) return nullptr;
    const int t = tile & 0xff;
    return (tile & SVPF_TILE_EXT) ? batched_kernel<true>(bs, t) : batched_kernel<false>(bs, t);
}
}  // namespace

// Workgroups of the given tile resident per CU for filters of up to nmax particles (0: unknown tile).
extern "C" int svpf_batched_wg_per_cu(int bs, int tile, int nmax) {
    if (svpf_kernels_init() != hipSuccess) return 0;
    const void* fn = batched_fn(bs, tile);
    if (!fn) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, bs, LdsLayout::bytes(nmax)) != hipSuccess) return 0;
    return nb;
}

extern "C" int svpf_tile_capacity(int tile, int bs) {
    if (bs != 256 && bs != 512 && bs != 1024) return 0;
    if (tile & ~(SVPF_TILE_EXT | 0xff)) return 0;
    tile &= 0xff;
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

extern "C" hipError_t svpf_launch_batched(int bs, int tile, int grid, const svpf_fparams* params,
                                          svpf_fstate* fstate, const svpf_fdesc* desc, float* pool_h,
                                          float* pool_hp, float* pool_v, const int* order, int nf, int* queue,
                                          int* progress, int nmax, int t0, int t1, hipStream_t stream) {
    const void* fn = batched_fn(bs, tile);
    if (!fn || grid <= 0) return hipErrorInvalidValue;
    size_t lds = LdsLayout::bytes(nmax);
    if (queue && !progress) return hipErrorInvalidValue;
    if (queue) svpf_queue_reset_kernel<<<1, 256, 0, stream>>>(queue, progress, nf);
    void* args[] = {&params, &fstate, &desc, &pool_h, &pool_hp, &pool_v, &order, &nf, &queue, &progress, &nmax,
                    &t0, &t1};
    hipError_t e = hipLaunchKernel(fn, dim3(grid), dim3(bs), args, lds, stream);
    if (e != hipSuccess) return e;
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
                                                       int stride, int bwg, int t0, int t1) {
    extern __shared__ __align__(16) char lds[];
    if ((int)blockIdx.x % stride) return;
    const int g = (int)blockIdx.x / stride;
    if (g > G || (g == G && !bwg)) return;  // g == G: the median workgroup
    const svpf_fdesc* D = desc;
    const svpf_fparams* P = &params[D->param_idx];
    CoopIO io{P, fstate, D, pool_h + D->off, pool_hp + D->off, pool_v + D->off, xbuf, G, g, bwg};
    CoopEngine<BS, CS> e;
    e.bind(lds, P->n, G, g);
    // a timed-out wait: the status word is set atomically and never written by the commit,
    // which no launch with a timed-out wait reaches
    if (!e.run(io, t0, t1) && threadIdx.x == 0) atomicOr(&fstate->status, 2);
}

#define SVPF_COOP_CASE(CSV)                                                                                  \
    case CSV:                                                                                                \
        fn = (const void*)svpf_coop_kernel<256, CSV>;                                                        \
        lds = CoopEngine<256, CSV>::lds_bytes(n);                                                            \
        break;

extern "C" int svpf_coop_rows(int cs) { return 256 / cs; }

#ifdef SVPF_TRACE
__device__ unsigned long long svpf_trace_buf[256 * 64];
__device__ unsigned svpf_trace_tick;
extern "C" hipError_t svpf_trace_dump(unsigned long long* out) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(svpf_trace_buf), sizeof(unsigned long long) * 256 * 64);
}
#endif

extern "C" size_t svpf_coop_lds_bytes(int n) { return CoopEngine<256, 1>::lds_bytes(n); }
extern "C" size_t svpf_coop_xbuf_words_host(int n) { return svpf_coop_xbuf_words(n); }

extern "C" hipError_t svpf_launch_coop(int cs, int G, int stride, const svpf_fparams* params, svpf_fstate* fstate,
                                       const svpf_fdesc* desc, float* pool_h, float* pool_hp, float* pool_v,
                                       uint64_t* xbuf, int n, int t0, int t1, int cooperative, int bwg,
                                       hipStream_t stream) {
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
    if (G < 1 || stride < 1 || (long)G * (256 / cs) < n || G * 8 > n || (bwg != 0 && bwg != 1))
        return hipErrorInvalidValue;
    if (lds > 159 * 1024) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    void* args[] = {(void*)&params, (void*)&fstate, (void*)&desc, (void*)&pool_h, (void*)&pool_hp, (void*)&pool_v,
                    (void*)&xbuf, (void*)&G, (void*)&stride, (void*)&bwg, (void*)&t0, (void*)&t1};
    dim3 grid((G + bwg) * stride), block(256);
    if (cooperative) return hipLaunchCooperativeKernel(fn, grid, block, args, (unsigned)lds, stream);
    return hipLaunchKernel(fn, grid, block, args, lds, stream);
}
