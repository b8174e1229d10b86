// svpf_device.hpp -- device helpers for the gfx950 SVPF kernels: the
// counter-based noise generator, wave64 / workgroup reductions, and the
// in-LDS sort + exact order-statistic selection used by the median bandwidth.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svpf_types.h"

#define SVPF_DEV __device__ __forceinline__

namespace svpf {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Counter = {particle, tick, purpose, filter}, key = seed.
// ---------------------------------------------------------------------------
SVPF_DEV void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                            uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Uniform on (0, 1] from 24 bits: exact in float32, so host and device agree bit for bit.
SVPF_DEV float u01_from_bits(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

// Box-Muller, cosine branch.
SVPF_DEV float normal_from_bits(uint32_t a, uint32_t b) {
    float u1 = u01_from_bits(a);
    float u2 = (float)(b >> 8) * (1.0f / 16777216.0f);
    return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}

struct CounterRng {
    uint32_t k0, k1, filter;
    SVPF_DEV void draw(uint32_t i, uint32_t tick, uint32_t purpose, uint32_t o[4]) const {
        philox4x32_10(i, tick, purpose, filter, k0, k1, o);
    }
    SVPF_DEV float normal(uint32_t i, uint32_t tick, uint32_t purpose) const {
        uint32_t o[4];
        draw(i, tick, purpose, o);
        return normal_from_bits(o[0], o[1]);
    }
    SVPF_DEV float uniform(uint32_t i, uint32_t tick, uint32_t purpose) const {
        uint32_t o[4];
        draw(i, tick, purpose, o);
        return u01_from_bits(o[2]);
    }
};

// ---------------------------------------------------------------------------
// Reductions (wave64 butterflies, then one LDS hop across waves).
// ---------------------------------------------------------------------------
SVPF_DEV float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
SVPF_DEV float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}
SVPF_DEV float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}
SVPF_DEV uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Reduce across the CS adjacent lanes that share one particle row.
template <int CS>
SVPF_DEV float group_sum(float v) {
#pragma unroll
    for (int o = CS / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
template <int CS>
SVPF_DEV float group_min(float v) {
#pragma unroll
    for (int o = CS / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}

// Workgroup reductions of up to 4 values.  `red` is LDS scratch of at least
// 4 * (blockDim / 64) floats.  Every thread receives the result.  Contains
// two barriers; all threads of the workgroup must call it.
template <int BS, int K>
SVPF_DEV void block_sum(float (&v)[K], float* red) {
    constexpr int NW = BS / kWave;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[k * NW + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < NW; ++q) s += red[k * NW + q];
        v[k] = s;
    }
    __syncthreads();
}

template <int BS>
SVPF_DEV float block_max(float v, float* red) {
    constexpr int NW = BS / kWave;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    v = wave_max(v);
    if (lane == 0) red[w] = v;
    __syncthreads();
    float m = red[0];
#pragma unroll
    for (int q = 1; q < NW; ++q) m = fmaxf(m, red[q]);
    __syncthreads();
    return m;
}

template <int BS>
SVPF_DEV float block_min(float v, float* red) {
    constexpr int NW = BS / kWave;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    v = wave_min(v);
    if (lane == 0) red[w] = v;
    __syncthreads();
    float m = red[0];
#pragma unroll
    for (int q = 1; q < NW; ++q) m = fminf(m, red[q]);
    __syncthreads();
    return m;
}

template <int BS>
SVPF_DEV uint32_t block_sum_u32(uint32_t v, uint32_t* red) {
    constexpr int NW = BS / kWave;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    v = wave_sum_u32(v);
    if (lane == 0) red[w] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) s += red[q];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// Bitonic sort of p2 floats in LDS (p2 a power of two), ascending.
// ---------------------------------------------------------------------------
template <int BS>
SVPF_DEV void bitonic_sort(float* a, int p2) {
    for (int k = 2; k <= p2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < (p2 >> 1); t += BS) {
                int i = 2 * j * (t / j) + (t % j);  // lower index of the pair
                int l = i + j;
                bool up = ((i & k) == 0);
                float x = a[i], y = a[l];
                if ((x > y) == up) {
                    a[i] = y;
                    a[l] = x;
                }
            }
            __syncthreads();
        }
    }
}

// Number of j in (i, n) with fl(fl(s[j]-s[i])^2) <= x, for ascending s.  The
// squared differences along a row are non-decreasing in j (rounding is
// monotone), so a binary search finds the boundary.
SVPF_DEV int row_count_le(const float* s, int i, int n, float x) {
    int lo = i + 1, hi = n;  // first index with r2 > x lies in [lo, hi]
    const float si = s[i];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        float d = s[mid] - si;
        float r2 = d * d;
        if (r2 <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo - (i + 1);
}

// Lower median (rank floor((M-1)/2), M = n(n-1)/2) of the pairwise squared
// distances of the ascending array s[0, n).  Bisection over the bit patterns
// of non-negative floats, which order like the unsigned integers: the result
// is the bit pattern of an actual pair value, identical to what a sort of all
// M float32 values selects.
template <int BS>
SVPF_DEV float median_pairwise_r2(const float* s, int n, uint32_t* red) {
    if (n < 2) return 0.0f;
    const uint64_t M = (uint64_t)n * (uint64_t)(n - 1) / 2;
    const uint32_t need = (uint32_t)((M - 1) / 2) + 1u;  // count(<= x) must reach rank + 1
    float dmax = s[n - 1] - s[0];
    uint32_t lo = 0u, hi = __float_as_uint(dmax * dmax);
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        float x = __uint_as_float(mid);
        uint32_t c = 0;
        for (int i = threadIdx.x; i < n - 1; i += BS) c += (uint32_t)row_count_le(s, i, n, x);
        c = block_sum_u32<BS>(c, red);
        if (c >= need) hi = mid;
        else lo = mid + 1u;
    }
    return __uint_as_float(hi);
}

SVPF_DEV float clamp_logvol(float h) { return fminf(fmaxf(h, -15.0f), 5.0f); }
SVPF_DEV float safe_expf(float x) { return expf(fminf(x, 20.0f)); }

}  // namespace svpf
