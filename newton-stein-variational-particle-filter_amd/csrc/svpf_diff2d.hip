// svpf_diff2d.hip -- kernels of the two-factor SVPF (see svpf_diff2d.hpp):
// forward with checkpoints (also the batched two-factor filter), and the
// adjoint, one workgroup per filter.
#include <hip/hip_runtime.h>

#include "svpf_diff2d.hpp"
#include "svpf_kernels.h"

using namespace svpf;

template <int BS, int R>
__global__ __launch_bounds__(BS) void svpf_d2_forward_kernel(svpf_d2config C, const svpf_d2desc* __restrict__ desc,
                                                             int B) {
    extern __shared__ __align__(16) char lds[];
    const int f = (int)blockIdx.x;
    if (f >= B) return;
    const svpf_d2desc& D = desc[f];
    D2Engine<BS, R> e;
    e.bind(lds, C.n, C.n);
    const D2Theta th = d2_theta(D.theta);
    const int n = C.n;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = e.row(r);
        e.xf[r] = i < n ? D.x0[i] : 0.f;
        e.xs[r] = i < n ? D.x0[n + i] : 0.f;
        e.vf[r] = e.vs[r] = 0.f;
    }
    float med[2] = {0.f, 0.f};
    const bool full_tape = C.C >= C.T;
    const size_t slotf = diff2d_slot_floats(n, C.S);
    for (int t = 0; t < C.T; ++t) {
        if (t % C.C == 0) {
            float* ck = D.ckpt + (size_t)(t / C.C) * 4 * n;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = e.row(r);
                if (i < n) {
                    ck[i] = e.xf[r];
                    ck[n + i] = e.xs[r];
                    ck[2 * n + i] = e.vf[r];
                    ck[3 * n + i] = e.vs[r];
                }
            }
        }
        float outv[8];
        if (full_tape) e.template tick<true, true>(C, th, D, t, D.tape + (size_t)t * slotf, med, outv);
        else e.template tick<false, true>(C, th, D, t, nullptr, med, outv);
        if (threadIdx.x == 0 && D.out) {
            for (int k = 0; k < SVPF_O2W; ++k) D.out[(size_t)t * SVPF_O2W + k] = outv[k];
        }
    }
    if (D.xT) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = e.row(r);
            if (i < n) {
                D.xT[i] = e.xf[r];
                D.xT[n + i] = e.xs[r];
            }
        }
    }
}

template <int BS, int R>
__global__ __launch_bounds__(BS) void svpf_d2_backward_kernel(svpf_d2config C, const svpf_d2desc* __restrict__ desc,
                                                              int B) {
    extern __shared__ __align__(16) char lds[];
    const int f = (int)blockIdx.x;
    if (f >= B) return;
    const svpf_d2desc& D = desc[f];
    D2Engine<BS, R> e;
    e.bind(lds, C.n, C.n);
    const D2Theta th = d2_theta(D.theta);
    const int n = C.n, T = C.T;
    const size_t slotf = diff2d_slot_floats(n, C.S);
    float xbf[R], xbs[R], vbf[R], vbs[R];
    double tb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) xbf[r] = xbs[r] = vbf[r] = vbs[r] = 0.f;
    const int nseg = (T + C.C - 1) / C.C;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int t0 = seg * C.C, t1 = min(T, t0 + C.C);
        if (C.C < T) {  // otherwise the forward wrote the tape
            const float* ck = D.ckpt + (size_t)seg * 4 * n;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = e.row(r);
                const bool ok = i < n;
                e.xf[r] = ok ? ck[i] : 0.f;
                e.xs[r] = ok ? ck[n + i] : 0.f;
                e.vf[r] = ok ? ck[2 * n + i] : 0.f;
                e.vs[r] = ok ? ck[3 * n + i] : 0.f;
            }
            float med[2] = {0.f, 0.f}, outv[8];
            for (int t = t0; t < t1; ++t)
                e.template tick<true, false>(C, th, D, t, D.tape + (size_t)(t - t0) * slotf, med, outv);
        }
        for (int t = t1 - 1; t >= t0; --t)
            e.tick_back(C, th, D, t, D.tape + (size_t)(t - t0) * slotf, D.gpl[t], xbf, xbs, vbf, vbs, tb);
    }
    constexpr int NW = BS / kWave;
    double* dr = reinterpret_cast<double*>(e.rec);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double s = tb[k];
        for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
        if (lane == 0) dr[k * NW + w] = s;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int q = 0; q < NW; ++q) s += dr[threadIdx.x * NW + q];
        D.gtheta[threadIdx.x] = (float)s;
    }
    if (D.gx0) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = e.row(r);
            if (i < n) {
                D.gx0[i] = xbf[r];
                D.gx0[n + i] = xbs[r];
            }
        }
    }
}

namespace {

bool d2_shape(int n, int* bs, int* rows) {
    if (n <= 256) *bs = 256, *rows = 1;
    else if (n <= 512) *bs = 512, *rows = 1;
    else if (n <= 1024) *bs = 1024, *rows = 1;
    else if (n <= 2048) *bs = 1024, *rows = 2;
    else return false;
    return true;
}

template <int BS, int R>
hipError_t launch_d2(bool backward, const svpf_d2config& C, const svpf_d2desc* desc, int B, hipStream_t stream) {
    const size_t lds = D2Lds::bytes(C.n);
    if (lds > 159 * 1024) return hipErrorInvalidValue;
    const void* fn = backward ? (const void*)svpf_d2_backward_kernel<BS, R> : (const void*)svpf_d2_forward_kernel<BS, R>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if (backward) svpf_d2_backward_kernel<BS, R><<<B, BS, lds, stream>>>(C, desc, B);
    else svpf_d2_forward_kernel<BS, R><<<B, BS, lds, stream>>>(C, desc, B);
    return hipGetLastError();
}

}  // namespace

extern "C" size_t svpf_diff2d_slot_floats_dev(int n, int S) { return diff2d_slot_floats(n, S); }

extern "C" hipError_t svpf_launch_diff2d(int backward, const svpf_d2config* C, const svpf_d2desc* desc, int B,
                                         hipStream_t stream) {
    int bs = 0, rows = 0;
    if (!C || C->n < 2 || C->S < 1 || C->T < 1 || C->C < 1 || B < 1 || !d2_shape(C->n, &bs, &rows))
        return hipErrorInvalidValue;
    const bool bw = backward != 0;
    if (bs == 256) return launch_d2<256, 1>(bw, *C, desc, B, stream);
    if (bs == 512) return launch_d2<512, 1>(bw, *C, desc, B, stream);
    if (rows == 1) return launch_d2<1024, 1>(bw, *C, desc, B, stream);
    return launch_d2<1024, 2>(bw, *C, desc, B, stream);
}
