// svpf_device.hpp -- device helpers for the gfx950 SVPF kernels: the
// counter-based noise generator, wave64 / workgroup reductions, and the
// in-LDS sort + exact order-statistic selection used by the median bandwidth.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svpf_types.h"

#define SVPF_DEV __device__ __forceinline__

// Phase timestamps for latency analysis (off unless built with -DSVPF_TRACE):
// thread 0 of block 0 stores the 100 MHz wall clock into slot k of the
// tick's 64-slot row.
// SVPF_TRACE_ROUND(k) stamps slot 51 + k of the tick last stamped at slot 0 (the
// rounds of the median selection).
#ifdef SVPF_TRACE
extern __device__ unsigned long long svpf_trace_buf[256 * 64];
extern __device__ unsigned svpf_trace_tick;
#define SVPF_TRACE_POINT(tick, k)                                                         \
    do {                                                                                  \
        if (threadIdx.x == 0 && blockIdx.x == 0) {                                        \
            svpf_trace_buf[(((unsigned)(tick)) & 255u) * 64u + ((k) & 63)] = wall_clock64(); \
            if ((k) == 0) svpf_trace_tick = (unsigned)(tick);                             \
        }                                                                                 \
    } while (0)
#define SVPF_TRACE_ROUND(k)                                                                              \
    do {                                                                                                 \
        if (threadIdx.x == 0 && blockIdx.x == 0 && (k) < 12)                                             \
            svpf_trace_buf[(svpf_trace_tick & 255u) * 64u + 51u + (unsigned)(k)] = wall_clock64();       \
    } while (0)
// thread 0 of the calling workgroup stamps slot k of the tick (for a workgroup other than 0)
#define SVPF_TRACE_STAMP(tick, k)                                                                \
    do {                                                                                         \
        if (threadIdx.x == 0) svpf_trace_buf[(((unsigned)(tick)) & 255u) * 64u + ((k) & 63)] = wall_clock64(); \
    } while (0)
#else
#define SVPF_TRACE_POINT(tick, k) \
    do {                          \
    } while (0)
#define SVPF_TRACE_ROUND(k) \
    do {                    \
    } while (0)
#define SVPF_TRACE_STAMP(tick, k) \
    do {                          \
    } while (0)
#endif

namespace svpf {

constexpr int kWave = 64;
constexpr int kMedianCands = 32;  // candidate pairs of the final rank selection by counting
constexpr int kHistBins = 256;    // bins of a histogram pass over the gathered candidates
constexpr int kHistCands = 1024;  // candidates gathered when the scratch holds the bins too
// thresholds counted per selection round (15 measured slower inside one workgroup:
// the wider lockstep searches cost more per round than the rounds they save)
constexpr int kMedianW = 7;
// LDS words of the reduction scratch `red`: the median selection uses 48 + BS/64 + 1
// (65 at 1024 threads), the block reductions fewer.
constexpr int kRedWords = 80;

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
// Reductions.  A wave64 reduction is a DPP inclusive scan -- row_shr 1/2/4/8
// inside each 16-lane row, then row_bcast 15 and 31 across rows -- after
// which lane 63 holds the total, broadcast with readlane.  Lanes whose DPP
// source lies outside the row keep the identity element.  All 64 lanes must
// be active.  Workgroup reductions add one LDS hop across waves.
// ---------------------------------------------------------------------------
#define SVPF_DPP_F(OLD, V, CTRL, RM) \
    __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(OLD), __float_as_int(V), CTRL, RM, 0xF, false))
#define SVPF_DPP_U(OLD, V, CTRL, RM) ((uint32_t)__builtin_amdgcn_update_dpp((int)(OLD), (int)(V), CTRL, RM, 0xF, false))

SVPF_DEV float wave_sum(float v) {
    v += SVPF_DPP_F(0.0f, v, 0x111, 0xF);  // row_shr:1
    v += SVPF_DPP_F(0.0f, v, 0x112, 0xF);  // row_shr:2
    v += SVPF_DPP_F(0.0f, v, 0x114, 0xF);  // row_shr:4
    v += SVPF_DPP_F(0.0f, v, 0x118, 0xF);  // row_shr:8
    v += SVPF_DPP_F(0.0f, v, 0x142, 0xA);  // row_bcast:15
    v += SVPF_DPP_F(0.0f, v, 0x143, 0xC);  // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
SVPF_DEV float wave_max(float v) {
    const float lo = -3.402823466e38f;
    v = fmaxf(v, SVPF_DPP_F(lo, v, 0x111, 0xF));
    v = fmaxf(v, SVPF_DPP_F(lo, v, 0x112, 0xF));
    v = fmaxf(v, SVPF_DPP_F(lo, v, 0x114, 0xF));
    v = fmaxf(v, SVPF_DPP_F(lo, v, 0x118, 0xF));
    v = fmaxf(v, SVPF_DPP_F(lo, v, 0x142, 0xA));
    v = fmaxf(v, SVPF_DPP_F(lo, v, 0x143, 0xC));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
SVPF_DEV float wave_min(float v) {
    const float hi = 3.402823466e38f;
    v = fminf(v, SVPF_DPP_F(hi, v, 0x111, 0xF));
    v = fminf(v, SVPF_DPP_F(hi, v, 0x112, 0xF));
    v = fminf(v, SVPF_DPP_F(hi, v, 0x114, 0xF));
    v = fminf(v, SVPF_DPP_F(hi, v, 0x118, 0xF));
    v = fminf(v, SVPF_DPP_F(hi, v, 0x142, 0xA));
    v = fminf(v, SVPF_DPP_F(hi, v, 0x143, 0xC));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
SVPF_DEV uint32_t wave_sum_u32(uint32_t v) {
    v += SVPF_DPP_U(0u, v, 0x111, 0xF);
    v += SVPF_DPP_U(0u, v, 0x112, 0xF);
    v += SVPF_DPP_U(0u, v, 0x114, 0xF);
    v += SVPF_DPP_U(0u, v, 0x118, 0xF);
    v += SVPF_DPP_U(0u, v, 0x142, 0xA);
    v += SVPF_DPP_U(0u, v, 0x143, 0xC);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// The value of lane (l ^ J) for J < 64: DPP for partners inside a 16-lane
// row (quad_perm for 1 and 2, row_shl / row_shr for 4 and 8), ds_bpermute
// across rows.  All lanes of the wave must be active.
template <int J>
SVPF_DEV float xor_lane(float x) {
    const int lane = threadIdx.x & (kWave - 1);
    if constexpr (J == 1) {
        return SVPF_DPP_F(x, x, 0xB1, 0xF);  // quad_perm [1,0,3,2]
    } else if constexpr (J == 2) {
        return SVPF_DPP_F(x, x, 0x4E, 0xF);  // quad_perm [2,3,0,1]
    } else if constexpr (J == 4) {
        float up = SVPF_DPP_F(x, x, 0x104, 0xF);  // row_shl:4, lane l reads l + 4
        float dn = SVPF_DPP_F(x, x, 0x114, 0xF);  // row_shr:4, lane l reads l - 4
        return (lane & 4) ? dn : up;
    } else if constexpr (J == 8) {
        float up = SVPF_DPP_F(x, x, 0x108, 0xF);
        float dn = SVPF_DPP_F(x, x, 0x118, 0xF);
        return (lane & 8) ? dn : up;
    } else {
        return __shfl_xor(x, J, kWave);
    }
}

// Reduce across the CS adjacent lanes that share one particle row (all lanes active).
template <int CS>
SVPF_DEV float group_sum(float v) {
    if constexpr (CS >= 2) v += xor_lane<1>(v);
    if constexpr (CS >= 4) v += xor_lane<2>(v);
    if constexpr (CS >= 8) v += xor_lane<4>(v);
    if constexpr (CS >= 16) v += xor_lane<8>(v);
    if constexpr (CS >= 32) v += xor_lane<16>(v);
    if constexpr (CS >= 64) v += xor_lane<32>(v);
    return v;
}
template <int CS>
SVPF_DEV float group_min(float v) {
    if constexpr (CS >= 2) v = fminf(v, xor_lane<1>(v));
    if constexpr (CS >= 4) v = fminf(v, xor_lane<2>(v));
    if constexpr (CS >= 8) v = fminf(v, xor_lane<4>(v));
    if constexpr (CS >= 16) v = fminf(v, xor_lane<8>(v));
    if constexpr (CS >= 32) v = fminf(v, xor_lane<16>(v));
    if constexpr (CS >= 64) v = fminf(v, xor_lane<32>(v));
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

// Inclusive prefix sum over the 64 lanes (the scan inside wave_sum_u32).
SVPF_DEV uint32_t wave_scan_u32(uint32_t v) {
    v += SVPF_DPP_U(0u, v, 0x111, 0xF);
    v += SVPF_DPP_U(0u, v, 0x112, 0xF);
    v += SVPF_DPP_U(0u, v, 0x114, 0xF);
    v += SVPF_DPP_U(0u, v, 0x118, 0xF);
    v += SVPF_DPP_U(0u, v, 0x142, 0xA);
    v += SVPF_DPP_U(0u, v, 0x143, 0xC);
    return v;
}

// ---------------------------------------------------------------------------
// Workgroup sort of n floats in LDS.
//
// Stage 1: every wave sorts runs of 128 values in registers (two per lane,
// element lane and lane + 64) with a bitonic network; partners at lane
// distance 1, 2 (quad_perm) and 4, 8 (row_shl / row_shr) come from DPP, 16
// and 32 from ds_bpermute, 64 is the other register.  Stage 2: runs are
// merged by rank -- an element's output index is its index in its own run
// plus the number of smaller elements in the other runs of its merge group
// (ties go to the lower run, so equal keys never collide) -- two runs at a
// time until at most four remain, then all four at once.
// ---------------------------------------------------------------------------
template <int K, int J>
SVPF_DEV void bitonic_step(float& x0, float& x1) {
    if constexpr (J == 64) {
        const float a = fminf(x0, x1), b = fmaxf(x0, x1);
        x0 = a;
        x1 = b;
    } else {
        const int lane = threadIdx.x & (kWave - 1);
        const float y0 = xor_lane<J>(x0), y1 = xor_lane<J>(x1);
        const bool lower = (lane & J) == 0;
        const bool up0 = (lane & K) == 0;
        const bool up1 = ((lane + 64) & K) == 0;
        x0 = (lower == up0) ? fminf(x0, y0) : fmaxf(x0, y0);
        x1 = (lower == up1) ? fminf(x1, y1) : fmaxf(x1, y1);
    }
}

template <int K, int J>
SVPF_DEV void bitonic_merge_steps(float& x0, float& x1) {
    bitonic_step<K, J>(x0, x1);
    if constexpr (J > 1) bitonic_merge_steps<K, J / 2>(x0, x1);
}

template <int K>
SVPF_DEV void bitonic_stages(float& x0, float& x1) {
    bitonic_merge_steps<K, K / 2>(x0, x1);
    if constexpr (K < 128) bitonic_stages<K * 2>(x0, x1);
}

// Ascending sort of the 128 values {x0 of lane l = element l, x1 = element l + 64}.
SVPF_DEV void wave_sort128(float& x0, float& x1) { bitonic_stages<2>(x0, x1); }

// One merge level over full runs: groups of WAYS runs of length lr (a power
// of two) from src[0, total) into dst, total a multiple of lr * WAYS and of
// BS.  An element's output index is its index within its group's own run
// plus the number of elements of every other run of the group that precede
// it in a stable merge (<= x in lower runs, < x in higher runs), so equal
// keys never collide.  Each thread ranks two elements at once; the binary
// searches advance in lockstep, branch-free, and never leave their run, so
// all loads of one step are in flight together.
template <int BS, int WAYS>
SVPF_DEV void merge_level(const float* src, float* dst, int total, int lr) {
    const int gl = lr * WAYS;
    const int lg = 31 - __clz(lr);
    for (int e0 = threadIdx.x; e0 < total; e0 += 2 * BS) {
        const bool two = e0 + BS < total;  // uniform: total is a multiple of BS
        float x[2];
        int out[2], own[2], pos[2][WAYS];
        bool strict[2][WAYS];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int e = two ? e0 + k * BS : e0;
            x[k] = src[e];
            const int gs = e & ~(gl - 1);
            const int q = (e - gs) >> lg;
            out[k] = e - q * lr;  // group start + index within the own run
            own[k] = q;
#pragma unroll
            for (int r = 0; r < WAYS; ++r) {
                pos[k][r] = gs + r * lr - 1;  // last index known to precede x (the own run is searched too, unused)
                strict[k][r] = r > q;
            }
        }
        // steps lr/2 .. 1 reach at most the run's second-to-last element; its
        // last element is tested after the loop
        for (int step = lr >> 1; step > 0; step >>= 1) {
            float u[2][WAYS];
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int r = 0; r < WAYS; ++r) u[k][r] = src[pos[k][r] + step];
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int r = 0; r < WAYS; ++r) {
                    const bool before = strict[k][r] ? (u[k][r] < x[k]) : (u[k][r] <= x[k]);
                    pos[k][r] += step & -(int)before;
                }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int gs = (two ? e0 + k * BS : e0) & ~(gl - 1);
#pragma unroll
            for (int r = 0; r < WAYS; ++r) {
                const int a = gs + r * lr;
                const float u = src[a + lr - 1];
                const bool before = strict[k][r] ? (u < x[k]) : (u <= x[k]);
                const int rank = before ? lr : pos[k][r] + 1 - a;
                out[k] += (r == own[k]) ? 0 : rank;
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
            if (k == 0 || two) dst[out[k]] = x[k];
    }
}

// Half-cleaner steps J = 32 .. 1 of an ascending merge inside the wave (partners
// at lane distance J; the lower lane keeps the minimum), for both registers.
template <int J>
SVPF_DEV void wave_clean_steps(float& x0, float& x1) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool lower = (lane & J) == 0;
    const float y0 = xor_lane<J>(x0), y1 = xor_lane<J>(x1);
    x0 = lower ? fminf(x0, y0) : fmaxf(x0, y0);
    x1 = lower ? fminf(x1, y1) : fmaxf(x1, y1);
    if constexpr (J > 1) wave_clean_steps<J / 2>(x0, x1);
}

// Stage 2 when the ps values fit the workgroup's registers (ps <= 2 BS, the
// 128-runs sorted ascending in x0 = element e0 = 128 w + lane, x1 = e0 + 64):
// bitonic merges in the form whose comparators all point one way -- for each
// block size K = 256 .. ps, a flip step against the mirrored partner e ^ (K - 1),
// then half-cleaners at e ^ J for J = K/4 .. 1 -- with the steps of J >= 128
// through LDS (one barrier each, alternating dst and src) and the rest in
// registers.  Writes the result to dst.
template <int BS>
SVPF_DEV void block_bitonic_tail(float x0, float x1, float* src, float* dst, int ps) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int e0 = 128 * w + lane, e1 = e0 + 64;
    const bool act = e0 < ps;  // uniform per wave: ps is a multiple of 128
    int stage = 0;
    // partner e ^ mask; the lower element keeps the minimum
    auto lds_step = [&](int mask) {
        float* b = (stage & 1) ? src : dst;  // the first write goes to dst: src may still be read
        ++stage;
        if (act) {
            b[e0] = x0;
            b[e1] = x1;
        }
        __syncthreads();
        if (act) {
            const int p0 = e0 ^ mask, p1 = e1 ^ mask;
            const float y0 = b[p0], y1 = b[p1];
            const bool lo0 = e0 < p0, lo1 = e1 < p1;
            x0 = lo0 ? fminf(x0, y0) : fmaxf(x0, y0);
            x1 = lo1 ? fminf(x1, y1) : fmaxf(x1, y1);
        }
    };
    for (int K = 256; K <= ps; K <<= 1) {
        lds_step(K - 1);  // flip
        for (int J = K / 4; J >= 128; J >>= 1) lds_step(J);
        if (act) {
            const float a = fminf(x0, x1), c = fmaxf(x0, x1);  // J = 64: e0 is the lower element
            x0 = a;
            x1 = c;
            wave_clean_steps<32>(x0, x1);
        }
    }
    // the last LDS step read from dst when its index was even: wait for every partner read
    if (((stage - 1) & 1) == 0) __syncthreads();
    if (act) {
        dst[e0] = x0;
        dst[e1] = x1;
    }
    __syncthreads();
}

// Sorts src[0, n) ascending into dst[0, n); src is overwritten.  Works over
// the power-of-two domain ps = max(128, pow2_ceil(n)) with +inf padding, so
// src and dst must hold ps floats when n > 128 (n floats otherwise); dst
// ends with the padding in [n, ps).  Contains barriers: all threads of the
// workgroup call it; src must be complete (synchronised) on entry, dst is
// complete on return.
template <int BS>
SVPF_DEV void block_sort(float* src, float* dst, int n) {
    constexpr int NW = BS / kWave;
    constexpr float kInf = 3.402823466e38f;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    int ps = 128;
    while (ps < n) ps <<= 1;
    const int runs = ps / 128;
    if (runs > 1 && ps <= 2 * BS) {
        // every value in registers: wave-sorted runs, then bitonic merges
        const int i0 = 128 * w + lane, i1 = i0 + 64;
        float x0 = kInf, x1 = kInf;
        if (i0 < ps) {
            x0 = i0 < n ? src[i0] : kInf;
            x1 = i1 < n ? src[i1] : kInf;
            wave_sort128(x0, x1);
        }
        block_bitonic_tail<BS>(x0, x1, src, dst, ps);
        return;
    }
    const int wlim = runs == 1 ? n : ps;  // one run: no padding written
    int levels = 0;
    for (int r = runs; r > 4; r >>= 1) ++levels;
    if (runs > 1) ++levels;
    float* a = (levels & 1) ? src : dst;  // stage 1 target, so the last level lands in dst
    for (int c = w; c < runs; c += NW) {
        const int i0 = c * 128 + lane, i1 = i0 + 64;
        float x0 = i0 < n ? src[i0] : kInf;
        float x1 = i1 < n ? src[i1] : kInf;
        wave_sort128(x0, x1);
        if (i0 < wlim) a[i0] = x0;
        if (i1 < wlim) a[i1] = x1;
    }
    __syncthreads();
    float* b = (a == src) ? dst : src;
    int lr = 128, r = runs;
    while (r > 4) {
        merge_level<BS, 2>(a, b, ps, lr);
        __syncthreads();
        float* t = a;
        a = b;
        b = t;
        lr *= 2;
        r >>= 1;
    }
    if (r == 4) {
        merge_level<BS, 4>(a, b, ps, lr);
        __syncthreads();
    } else if (r == 2) {
        merge_level<BS, 2>(a, b, ps, lr);
        __syncthreads();
    }
}

// First index j in [lo, hi] with fl(fl(s[j]-si)^2) > x, for ascending s and a
// row whose squared differences are non-decreasing in j (rounding is
// monotone), given that the answer lies in [lo, hi].
SVPF_DEV int row_first_gt(const float* s, float si, int lo, int hi, float x) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        float d = s[mid] - si;
        if (d * d <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The W lockstep searches of one row: res[q] = first j in [a, b] with
// fl(fl(s[j]-si)^2) > xb[q].  Every j >= b fails the test (b ends the row's
// bracket, and s[n] is +inf), so probes need no bracket check: they are
// only clamped to n.
template <int W>
SVPF_DEV void row_search(const float* s, int n, float si, int a, int b, const uint32_t (&xb)[W], int (&res)[W]) {
    int pos[W];
#pragma unroll
    for (int q = 0; q < W; ++q) pos[q] = a - 1;
    const int len = b - a;
    for (int step = len > 0 ? (1 << (31 - __clz(len))) : 0; step > 0; step >>= 1) {
        float u[W];
#pragma unroll
        for (int q = 0; q < W; ++q) u[q] = s[min(pos[q] + step, n)];  // all W loads in flight
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const float d = u[q] - si;
            pos[q] += step & -(int)(d * d <= __uint_as_float(xb[q]));
        }
    }
#pragma unroll
    for (int q = 0; q < W; ++q) res[q] = pos[q] + 1;
}

// The value of rank need - 1 among M values, given C of them in cand[0, C) (LDS),
// exactly those whose bit patterns lie in [L, H], with cL values below L.  With
// bins (kHistBins words of LDS, free for this call; cand must not overlap them),
// histogram passes -- kHistBins equal bins of bit patterns over [L, H], LDS
// atomics, one prefix scan per wave -- narrow [L, H] to the bin holding the rank
// until at most kMedianCands candidates remain, which are compacted into the
// bins' words; the rank is then read off by counting (without bins C must be at
// most kMedianCands).  ctr, res: one word of LDS each.  All threads call it.
template <int BS>
SVPF_DEV float select_candidates(const float* cand, int C, uint32_t L, uint32_t H, uint32_t cL, uint32_t need,
                                 uint32_t* hist, uint32_t* ctr, uint32_t* res) {
    const int lane = threadIdx.x & (kWave - 1);
    if (L >= H) return __uint_as_float(H);
    uint32_t cH = cL + (uint32_t)C;
    if (hist && C > kMedianCands) {
        const int C0 = C;
        bool first = true;
        while (cH - cL > (uint32_t)kMedianCands && L < H) {
            const uint32_t span = H - L;  // bins cover [L, L + (kHistBins << sh)) >= [L, H]
            int sh = 0;
            while (((uint64_t)kHistBins << sh) <= (uint64_t)span) ++sh;
            if (!first) __syncthreads();  // every wave has read the previous pass's bins
            first = false;
            for (int k = threadIdx.x; k < kHistBins; k += BS) hist[k] = 0u;
            __syncthreads();
            for (int k = threadIdx.x; k < C0; k += BS) {
                const uint32_t x = __float_as_uint(cand[k]);
                if (x >= L && x <= H) atomicAdd(&hist[(x - L) >> sh], 1u);
            }
            __syncthreads();
            // every wave scans the bins (identical results): lane l holds bins 4l .. 4l+3
            static_assert(kHistBins == 4 * kWave, "four bins per lane");
            const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                           h3 = hist[4 * lane + 3];
            const uint32_t incl = wave_scan_u32(h0 + h1 + h2 + h3);
            const uint64_t hit = __ballot(cL + incl >= need);
            // the bins hold every candidate of [L, H], so some lane reaches the rank
            const int ls = __builtin_amdgcn_readfirstlane(hit ? (int)__builtin_ctzll(hit) : kWave - 1);
            uint32_t base = cL + (uint32_t)__builtin_amdgcn_readlane((int)(incl - (h0 + h1 + h2 + h3)), ls);
            const uint32_t g[4] = {(uint32_t)__builtin_amdgcn_readlane((int)h0, ls),
                                   (uint32_t)__builtin_amdgcn_readlane((int)h1, ls),
                                   (uint32_t)__builtin_amdgcn_readlane((int)h2, ls),
                                   (uint32_t)__builtin_amdgcn_readlane((int)h3, ls)};
            int b = 0;
            while (b < 3 && base + g[b] < need) base += g[b++];
            const uint32_t nL = L + ((uint32_t)(4 * ls + b) << sh);
            const uint32_t top = nL + ((1u << sh) - 1u);
            H = top < H ? top : H;
            L = nL;
            cL = base;
            cH = base + g[b];
        }
        if (L >= H) return __uint_as_float(H);
        // compact the candidates left in [L, H] into the bins' words
        __syncthreads();
        if (threadIdx.x == 0) *ctr = 0u;
        __syncthreads();
        float* small = reinterpret_cast<float*>(hist);
        for (int k = threadIdx.x; k < C0; k += BS) {
            const float v = cand[k];
            const uint32_t x = __float_as_uint(v);
            if (x >= L && x <= H) {
                const uint32_t q = atomicAdd(ctr, 1u);
                if (q < (uint32_t)kHistBins) small[q] = v;
            }
        }
        __syncthreads();
        cand = small;
        C = (int)(cH - cL);
    }
    const uint32_t r0 = need - cL - 1u;
    for (int k = threadIdx.x; k < C; k += BS) {
        const float v = cand[k];
        uint32_t below = 0u, upto = 0u;
        for (int m = 0; m < C; ++m) {
            const float u = cand[m];
            below += (u < v) ? 1u : 0u;
            upto += (u <= v) ? 1u : 0u;
        }
        if (below <= r0 && r0 < upto) *res = __float_as_uint(v);
    }
    __syncthreads();
    return __uint_as_float(*res);
}

// Lower median (rank floor((M-1)/2), M = n(n-1)/2) of the pairwise squared
// distances of the ascending array s[0, n), bit-identical to sorting all M
// float32 values.  s needs room for n + 1 floats: s[n] is set to +inf.
//
// Multi-way bisection over the bit patterns of non-negative floats (which
// order like unsigned integers, and whose spacing is close to geometric):
// each round counts, with W lockstep binary searches per row, the pairs at
// or below W thresholds inside the bracket [L, H], keeps the sub-bracket
// holding the rank, and narrows every row's index bracket [plo[i], phi[i])
// to the pairs still inside.  The thresholds of the first round are spaced
// 1/16 octave around `est` (the previous tick's value; when <= 0, half the
// squared interquartile range, which the median of r^2 equals for normal
// data); while the rank stays below (above) all of them the next round
// extends the grid downwards (upwards) with steps 8x wider, and once it is
// bracketed the bracket is split evenly.  Once few enough candidate pairs
// remain they are gathered: with cap >= kHistBins + 64 words of scratch at scr,
// up to kHistCands, which histogram passes (kHistBins equal bins of bit
// patterns over the bracket, LDS atomics, one prefix scan) narrow to at most
// kMedianCands -- a pass over a thousand gathered values costs less than a
// round of binary searches over every row; with less scratch, kMedianCands.
// The rank is then read off by counting.  Scratch (LDS): plo, phi (n Idx
// each), scr (cap words), red (48 + BS/64 + 1 words); Idx may be a 16-bit type
// when n < 65536.  The first KR row slots of each thread keep their search
// results in registers; later slots search again.
template <int BS, int KR, typename Idx = int, int W = kMedianW>
SVPF_DEV float median_pairwise_r2(float* s, int n, float est, Idx* plo, Idx* phi, float* scr, int cap,
                                  uint32_t* red) {
    static_assert(W <= 16, "counts: three rotating 16-word buffers");
    constexpr int NW = BS / kWave;
    constexpr uint32_t kSeedStep = 1u << 19;  // 1/16 octave of r^2 (in float bit patterns)
    enum { SEED = 0, DOWN = 1, UP = 2, SPLIT = 3 };
    constexpr int kScan = 48, kResult = 48 + NW;
    static_assert(kResult < 80, "reduction scratch: red holds 48 + BS/64 + 1 words");
    if (n < 2) return 0.0f;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int rows = n - 1;
    const uint32_t M = (uint32_t)(((uint64_t)n * (uint64_t)(n - 1)) / 2);
    const uint32_t need = (M - 1u) / 2u + 1u;  // count(<= x) must reach rank + 1
    // scratch: with room for the bins, up to kHistCands candidates are gathered and
    // narrowed by histogram passes; otherwise kMedianCands, ranked by counting
    const bool hsel = cap >= kHistBins + 2 * kMedianCands;
    uint32_t* hist = reinterpret_cast<uint32_t*>(scr);
    float* cand = hsel ? scr + kHistBins : scr;
    const int cmax = hsel ? min(cap - kHistBins, kHistCands) : min(cap, kMedianCands);
    const float dmax = s[n - 1] - s[0];
    uint32_t L = 0u, H = __float_as_uint(dmax * dmax);
    uint32_t cL = 0u, cH = M;  // pairs below L, pairs at or below H
    for (int i = threadIdx.x; i < rows; i += BS) {
        plo[i] = i + 1;
        phi[i] = n;
    }
    if (threadIdx.x < 16) red[threadIdx.x] = 0u;
    if (threadIdx.x == 0) s[n] = 3.402823466e38f;
    __syncthreads();
    if (!(est > 0.0f && est < 3.0e38f)) {
        const float iq = s[(3 * n) / 4] - s[n / 4];
        est = 0.5f * iq * iq;
    }
    const uint32_t eb = (est > 0.0f && est < 3.0e38f) ? __float_as_uint(est) : 0u;
    int mode = (eb > L && eb < H) ? SEED : SPLIT;
    int64_t anchor = (int64_t)eb, step = (int64_t)kSeedStep;
    int round = 0;
    while (cH - cL > (uint32_t)cmax && L < H) {
        // thresholds, non-decreasing in q, inside [L, H)
        auto thr = [&](int q) -> uint32_t {
            int64_t x;
            if (mode == SEED) x = anchor + (int64_t)(q - W / 2) * step;
            else if (mode == DOWN) x = anchor - (int64_t)(W - q) * step;
            else if (mode == UP) x = anchor + (int64_t)(q + 1) * step;
            else return L + (uint32_t)(((uint64_t)(H - L) * (uint64_t)(q + 1)) / (W + 1));
            x = x < (int64_t)L ? (int64_t)L : x;
            x = x > (int64_t)H - 1 ? (int64_t)H - 1 : x;
            return (uint32_t)x;
        };
        uint32_t xb[W];
#pragma unroll
        for (int q = 0; q < W; ++q) xb[q] = thr(q);
        uint32_t c[W];
#pragma unroll
        for (int q = 0; q < W; ++q) c[q] = 0u;
        int keep[KR][W];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int i = threadIdx.x + k * BS;
            if (i < rows) {
                row_search<W>(s, n, s[i], plo[i], phi[i], xb, keep[k]);
#pragma unroll
                for (int q = 0; q < W; ++q) c[q] += (uint32_t)(keep[k][q] - (i + 1));
            }
        }
        for (int i = threadIdx.x + KR * BS; i < rows; i += BS) {
            int res[W];
            row_search<W>(s, n, s[i], plo[i], phi[i], xb, res);
#pragma unroll
            for (int q = 0; q < W; ++q) c[q] += (uint32_t)(res[q] - (i + 1));
        }
        // counts: three rotating 16-word buffers, one barrier per round
        uint32_t* buf = red + 16 * (round % 3);
#pragma unroll
        for (int q = 0; q < W; ++q) c[q] = wave_sum_u32(c[q]);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < W; ++q) atomicAdd(&buf[q], c[q]);
        }
        if (threadIdx.x < 16) red[16 * ((round + 1) % 3) + threadIdx.x] = 0u;
        __syncthreads();
        int qs = W;  // first threshold reaching the rank (W: none, the upper bracket end)
#pragma unroll
        for (int q = W - 1; q >= 0; --q)
            if (buf[q] >= need) qs = q;
        qs = __builtin_amdgcn_readfirstlane(qs);
        const uint32_t xlo = qs > 0 ? thr(qs - 1) : 0u, xhi = qs < W ? thr(qs) : 0u;
        const uint32_t clo = qs > 0 ? buf[qs - 1] : 0u, chi = qs < W ? buf[qs] : 0u;
        if ((mode == SEED || mode == DOWN) && qs == 0) {
            anchor = (int64_t)thr(0);  // still below the grid: extend it downwards
            step *= 8;
            mode = DOWN;
        } else if ((mode == SEED || mode == UP) && qs == W) {
            anchor = (int64_t)thr(W - 1);
            step *= 8;
            mode = UP;
        } else {
            mode = SPLIT;
        }
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int i = threadIdx.x + k * BS;
            if (i < rows) {
                int na = plo[i], nb = phi[i];
#pragma unroll
                for (int q = 0; q < W; ++q) {  // bit blends, not selects: the results stay in registers
                    const int ma = -(int)(q == qs - 1), mb = -(int)(q == qs);
                    na = (keep[k][q] & ma) | (na & ~ma);
                    nb = (keep[k][q] & mb) | (nb & ~mb);
                }
                plo[i] = na;
                phi[i] = nb;
            }
        }
        for (int i = threadIdx.x + KR * BS; i < rows; i += BS) {
            const float si = s[i];
            int na = plo[i], nb = phi[i];
            if (qs > 0) na = row_first_gt(s, si, na, nb, __uint_as_float(xlo));
            if (qs < W) nb = row_first_gt(s, si, na, nb, __uint_as_float(xhi));
            plo[i] = na;
            phi[i] = nb;
        }
        if (qs > 0) {
            L = xlo + 1u;
            cL = clo;
        }
        if (qs < W) {
            H = xhi;
            cH = chi;
        }
        ++round;
        SVPF_TRACE_ROUND(round);
    }
    SVPF_TRACE_ROUND(11);
    if (L >= H) return __uint_as_float(H);
    // gather the cH - cL candidates (pairs with r^2 in [L, H]) and select by rank
    uint32_t cnt = 0u;
    for (int i = threadIdx.x; i < rows; i += BS) cnt += (uint32_t)(phi[i] - plo[i]);
    const uint32_t incl = wave_scan_u32(cnt);
    if (lane == kWave - 1) red[kScan + w] = incl;
    __syncthreads();
    uint32_t off = incl - cnt;
#pragma unroll
    for (int q = 0; q < NW; ++q) off += (q < w) ? red[kScan + q] : 0u;
    for (int i = threadIdx.x; i < rows; i += BS) {
        const float si = s[i];
        for (int j = plo[i]; j < phi[i]; ++j) {
            const float d = s[j] - si;
            cand[off++] = d * d;
        }
    }
    __syncthreads();
    const int C = (int)(cH - cL);
    return select_candidates<BS>(cand, C, L, H, cL, need, hsel ? hist : nullptr, red, red + kResult);
}

// First pair (a, b), a < b, in row-major order (least a * n + b) of the n values vals[] whose
// squared difference fl(fl(v_a - v_b)^2) equals r2 > 0 exactly; a = b = -1 when there is none.
// x holds the calling thread's rows (row threadIdx.x + r BS), vals the same values by index and
// sorted them ascending (both LDS, complete on entry).  On either side of a value x the squared
// difference fl((x - v)^2) is monotone in v, so two binary searches over sorted tell whether x has
// a partner; the rows that do are listed (list: n ints of LDS) and only pairs among them are
// checked.  The result equals a scan of every pair, in O(n log n + m^2) for m listed rows.
// words: 2 ints of LDS.  Every thread of the workgroup calls it.
template <int BS, int R>
SVPF_DEV void first_pair_at_r2(const float (&x)[R], const float* vals, const float* sorted, int n, float r2, int* list,
                               int* words, int& a, int& b) {
    if (threadIdx.x == 0) {
        words[0] = 0;
        words[1] = 0x7fffffff;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = (int)threadIdx.x + r * BS;
        if (i >= n) continue;
        const float xv = x[r];
        // right side: the first sorted value above xv whose squared difference reaches r2
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const float v = sorted[mid];
            const float d = xv - v;
            if (v > xv && d * d >= r2) hi = mid;
            else lo = mid + 1;
        }
        bool hit = false;
        if (lo < n) {
            const float d = xv - sorted[lo];
            hit = d * d == r2;
        }
        // left side: the last sorted value below xv whose squared difference reaches r2
        lo = 0;
        hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const float v = sorted[mid];
            const float d = xv - v;
            if (v < xv && d * d >= r2) lo = mid + 1;
            else hi = mid;
        }
        if (lo > 0) {
            const float d = xv - sorted[lo - 1];
            hit = hit || d * d == r2;
        }
        if (hit) list[atomicAdd(&words[0], 1)] = i;
    }
    __syncthreads();
    const int m = words[0];
    int best = 0x7fffffff;
    for (int p = (int)threadIdx.x; p < m * m; p += BS) {
        const int u = p / m, w = p - u * m;
        const int i = list[u], j = list[w];
        if (i < j) {
            const float d = vals[i] - vals[j];
            if (d * d == r2) best = min(best, i * n + j);
        }
    }
    if (best != 0x7fffffff) atomicMin(&words[1], best);
    __syncthreads();
    const int k = words[1];
    __syncthreads();
    if (k == 0x7fffffff) {
        a = b = -1;
    } else {
        a = k / n;
        b = k - a * n;
    }
}

SVPF_DEV float clamp_logvol(float h) { return fminf(fmaxf(h, -15.0f), 5.0f); }
SVPF_DEV float safe_expf(float x) { return expf(fminf(x, 20.0f)); }

}  // namespace svpf
