// svpf_2f.hip -- gfx950 kernels of the coordinate-wise two-factor step
// (svpf_2f.hpp): one workgroup per filter advancing ticks [t0, t1), and the
// initial draw of the factor particles.
#include <hip/hip_runtime.h>

#include "svpf_2f.hpp"
#include "svpf_kernels.h"

using namespace svpf;

template <int BS, int R, int CS>
__global__ __launch_bounds__(BS) void svpf_2f_kernel(const svpf_fparams* __restrict__ P, svpf_fstate* __restrict__ S,
                                                     const svpf_fdesc* __restrict__ D, float* __restrict__ x2f,
                                                     float* __restrict__ h, int t0, int t1) {
    extern __shared__ __align__(16) char lds[];
    Engine2F<BS, R, CS> e;
    e.bind(lds, P->n, P->n);
    typename Engine2F<BS, R, CS>::IO io{P, S, D, x2f, h};
    e.load2(io);
    for (int t = t0; t < t1; ++t) e.tick(io, t);
    e.store2(io);
}

// src/svpf_two_factor.cu:29-48: each factor from its stationary law, clamped.
__global__ void svpf_2f_init_kernel(float* x2f, int n, float sd_f, float sd_s, uint32_t k0, uint32_t k1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o[4];
    philox4x32_10((uint32_t)i, 0u, SVPF_PURPOSE_INIT_2F, 0u, k0, k1, o);
    const float f = clamp_logvol(sd_f * normal_from_bits(o[0], o[1]));
    const float s = clamp_logvol(sd_s * normal_from_bits(o[2], o[3]));
    x2f[i] = f;
    x2f[n + i] = s;
    x2f[2 * n + i] = f;
    x2f[3 * n + i] = s;
    x2f[4 * n + i] = 0.0f;
    x2f[5 * n + i] = 0.0f;
}

namespace {
const void* kernel_2f(int bs, int tile) {
#define SVPF_K2(BSV, RV, CSV) return (const void*)svpf_2f_kernel<BSV, RV, CSV>
    if (bs == 256) {
        switch (tile) {
            case SVPF_TILE_R1_C4: SVPF_K2(256, 1, 4);
            case SVPF_TILE_R1_C2: SVPF_K2(256, 1, 2);
            case SVPF_TILE_R1_C1: SVPF_K2(256, 1, 1);
            case SVPF_TILE_R2_C1: SVPF_K2(256, 2, 1);
            case SVPF_TILE_R4_C1: SVPF_K2(256, 4, 1);
            case SVPF_TILE_R8_C1: SVPF_K2(256, 8, 1);
            default: return nullptr;
        }
    }
    if (bs == 512 && tile == SVPF_TILE_R8_C1) SVPF_K2(512, 8, 1);
    return nullptr;
#undef SVPF_K2
}
}  // namespace

extern "C" hipError_t svpf_launch_2f(int bs, int tile, const svpf_fparams* P, svpf_fstate* S, const svpf_fdesc* D,
                                     float* x2f, float* h, int n, int t0, int t1, hipStream_t stream) {
    const void* fn = kernel_2f(bs, tile & 0xff);
    if (!fn || n < 1 || n > svpf_tile_capacity(tile & 0xff, bs)) return hipErrorInvalidValue;
    const size_t lds = LdsLayout::bytes(n);
    if (lds > 159 * 1024) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024);
    if (e != hipSuccess) return e;
    void* args[] = {(void*)&P, (void*)&S, (void*)&D, (void*)&x2f, (void*)&h, (void*)&t0, (void*)&t1};
    e = hipLaunchKernel(fn, dim3(1), dim3(bs), args, lds, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" hipError_t svpf_launch_2f_init(float* x2f, int n, float rho_f, float sigma_f, float rho_s, float sigma_s,
                                          uint64_t seed, hipStream_t stream) {
    if (n <= 0) return hipErrorInvalidValue;
    const float sd_f = sigma_f / sqrtf(1.0f - rho_f * rho_f + 1e-8f);
    const float sd_s = sigma_s / sqrtf(1.0f - rho_s * rho_s + 1e-8f);
    svpf_2f_init_kernel<<<(n + 255) / 256, 256, 0, stream>>>(x2f, n, sd_f, sd_s, (uint32_t)seed,
                                                             (uint32_t)(seed >> 32));
    return hipGetLastError();
}
