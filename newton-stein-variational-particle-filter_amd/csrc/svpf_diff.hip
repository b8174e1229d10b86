// svpf_diff.hip -- kernels of the differentiable SVPF (see svpf_diff.hpp):
// forward with checkpoints, and the adjoint, one workgroup per filter.
#include <hip/hip_runtime.h>

#include "svpf_diff.hpp"
#include "svpf_kernels.h"

using namespace svpf;

template <int BS, int R>
__global__ __launch_bounds__(BS) void svpf_diff_forward_kernel(svpf_dconfig C, const svpf_ddesc* __restrict__ desc,
                                                               int B) {
    extern __shared__ __align__(16) char lds[];
    const int f = (int)blockIdx.x;
    if (f >= B) return;
    const svpf_ddesc& D = desc[f];
    DiffEngine<BS, R> e;
    e.bind(lds, C.n, C.n);
    const DiffTheta th = diff_theta(D.theta);
    const int n = C.n;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = e.row(r);
        e.x[r] = i < n ? D.h0[i] : 0.f;
        e.v[r] = 0.f;
    }
    float med = 0.f, y_prev = 0.f;
    const bool full_tape = C.C >= C.T;
    const size_t slotf = diff_slot_floats(n, C.S);
    for (int t = 0; t < C.T; ++t) {
        if (t % C.C == 0) {
            float* ck = D.ckpt + (size_t)(t / C.C) * 2 * n;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = e.row(r);
                if (i < n) {
                    ck[i] = e.x[r];
                    ck[n + i] = e.v[r];
                }
            }
        }
        // checkpoint interval >= T: the forward writes the whole tape and the backward skips the recompute
        const float pl = full_tape ? e.template tick<true, true>(C, th, D, t, y_prev, D.tape + (size_t)t * slotf, med)
                                   : e.template tick<false, true>(C, th, D, t, y_prev, nullptr, med);
        if (threadIdx.x == 0) D.pred_ll[t] = pl;
        y_prev = D.y[t];
    }
    if (D.hT) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (e.row(r) < n) D.hT[e.row(r)] = e.x[r];
    }
}

template <int BS, int R>
__global__ __launch_bounds__(BS) void svpf_diff_backward_kernel(svpf_dconfig C, const svpf_ddesc* __restrict__ desc,
                                                                int B) {
    extern __shared__ __align__(16) char lds[];
    const int f = (int)blockIdx.x;
    if (f >= B) return;
    const svpf_ddesc& D = desc[f];
    DiffEngine<BS, R> e;
    e.bind(lds, C.n, C.n);
    const DiffTheta th = diff_theta(D.theta);
    const int n = C.n, T = C.T;
    const size_t slotf = diff_slot_floats(n, C.S);
    float xb[R], vb[R];
    double tb[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) xb[r] = vb[r] = 0.f;
    const int nseg = (T + C.C - 1) / C.C;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int t0 = seg * C.C, t1 = min(T, t0 + C.C);
        const float* ck = D.ckpt + (size_t)seg * 2 * n;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = e.row(r);
            e.x[r] = i < n ? ck[i] : 0.f;
            e.v[r] = i < n ? ck[n + i] : 0.f;
        }
        float med = 0.f, y_prev = t0 > 0 ? D.y[t0 - 1] : 0.f;
        if (C.C < T) {  // otherwise the forward wrote the tape
            for (int t = t0; t < t1; ++t) {
                e.template tick<true, false>(C, th, D, t, y_prev, D.tape + (size_t)(t - t0) * slotf, med);
                y_prev = D.y[t];
            }
        }
        for (int t = t1 - 1; t >= t0; --t) {
            const float yp = t > 0 ? D.y[t - 1] : 0.f;
            e.tick_back(C, th, D, t, yp, D.tape + (size_t)(t - t0) * slotf, D.gpl[t], xb, vb, tb);
        }
    }
    // dL/dtheta: wave sums, then one LDS hop (the transport records are free now)
    constexpr int NW = BS / kWave;
    double* dr = reinterpret_cast<double*>(e.rec);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double s = tb[k];
        for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
        if (lane == 0) dr[k * NW + w] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int q = 0; q < NW; ++q) s += dr[threadIdx.x * NW + q];
        D.gtheta[threadIdx.x] = (float)s;
    }
    if (D.gh0) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (e.row(r) < n) D.gh0[e.row(r)] = xb[r];
    }
}

namespace {

bool diff_shape(int n, int* bs, int* rows) {
    if (n <= 256) *bs = 256, *rows = 1;
    else if (n <= 512) *bs = 512, *rows = 1;
    else if (n <= 1024) *bs = 1024, *rows = 1;
    else if (n <= 2048) *bs = 1024, *rows = 2;
    else return false;
    return true;
}

template <int BS, int R>
hipError_t launch_diff(bool backward, const svpf_dconfig& C, const svpf_ddesc* desc, int B, hipStream_t stream) {
    const size_t lds = DiffLds::bytes(C.n);
    const void* fn = backward ? (const void*)svpf_diff_backward_kernel<BS, R> : (const void*)svpf_diff_forward_kernel<BS, R>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if (backward) svpf_diff_backward_kernel<BS, R><<<B, BS, lds, stream>>>(C, desc, B);
    else svpf_diff_forward_kernel<BS, R><<<B, BS, lds, stream>>>(C, desc, B);
    return hipGetLastError();
}

}  // namespace

extern "C" size_t svpf_diff_slot_floats_dev(int n, int S) { return diff_slot_floats(n, S); }

extern "C" hipError_t svpf_launch_diff(int backward, const svpf_dconfig* C, const svpf_ddesc* desc, int B,
                                       hipStream_t stream) {
    int bs = 0, rows = 0;
    if (!C || C->n < 2 || C->S < 1 || C->T < 1 || C->C < 1 || B < 1 || !diff_shape(C->n, &bs, &rows))
        return hipErrorInvalidValue;
    const bool bw = backward != 0;
    if (bs == 256) return launch_diff<256, 1>(bw, *C, desc, B, stream);
    if (bs == 512) return launch_diff<512, 1>(bw, *C, desc, B, stream);
    if (rows == 1) return launch_diff<1024, 1>(bw, *C, desc, B, stream);
    return launch_diff<1024, 2>(bw, *C, desc, B, stream);
}
