// Diagnostics beside the step: one-workgroup reductions over a filter's
// particle arrays, read after svpf_step_graph.
//
//   SVPF_DIAG_HEALTH  src/svpf_diagnostics.cu:126-221 (svpf_diag_check_kernel):
//                     NaN / Inf counts, the weight sums behind the ESS and the
//                     particle sum.
//   SVPF_DIAG_NU      src/tools/svpf_gradient_diagnostic.cu:61-180: the
//                     posterior-weighted mean of d/dnu log p(y_t | h, nu) and of z^2.
//   SVPF_DIAG_SIGMA   src/tools/svpf_gradient_diagnostic.cu:421-532: the
//                     posterior-weighted mean of d/dsigma log p(h_pred | h_prev)
//                     and of eps^2 / sigma^2.
//
// The work is O(N) with N <= 16384 floats per array (at most 192 KB read), so
// one workgroup of 1024 lanes does it in a few microseconds: the launch is the
// cost, not HBM.  Every reduction runs in a fixed order (strided lane sums,
// a butterfly inside each wave, waves summed in index order), so a result is
// bitwise reproducible.  The reference's ESS shift reduces the log-weight
// maximum over its first warp only; ESS does not depend on the shift, and the
// full maximum used here keeps every weight <= 1.
#include <hip/hip_runtime.h>

#include "svpf_kernels.h"
#include "svpf_special.h"

namespace {

constexpr int kDiagBS = 1024;
constexpr int kWaves = kDiagBS / 64;

struct SumOp {
    __device__ float operator()(float a, float b) const { return a + b; }
};
struct MaxOp {
    __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct MinOp {
    __device__ float operator()(float a, float b) const { return fminf(a, b); }
};

// All lanes return the workgroup's reduction; red[] holds kWaves partials.
template <class Op>
__device__ float block_reduce(float x, float* red, Op op) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = op(x, __shfl_xor(x, o));
    __syncthreads();  // red[] free (a previous reduction has been read)
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    float r = red[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) r = op(r, red[w]);
    return r;
}

__device__ bool finite_h(float h) { return h == h && h > -1e10f && h < 1e10f; }

__global__ __launch_bounds__(kDiagBS) void svpf_diag_kernel(int mode, const float* __restrict__ h,
                                                          const float* __restrict__ h2,
                                                          const float* __restrict__ lw, int n, float a, float b,
                                                          float c, float* __restrict__ out0,
                                                          float* __restrict__ out1) {
    __shared__ float red[kWaves];
    const int tid = threadIdx.x;
    if (mode == SVPF_DIAG_HEALTH) {
        float mx = -1e10f;
        for (int i = tid; i < n; i += kDiagBS) {
            const float l = lw[i];
            if (l == l && l < 1e10f) mx = fmaxf(mx, l);
        }
        mx = block_reduce(mx, red, MaxOp{});
        float nan_c = 0.f, inf_c = 0.f, sw = 0.f, sw2 = 0.f, sh = 0.f, hmin = 1e10f, hmax = -1e10f;
        for (int i = tid; i < n; i += kDiagBS) {
            const float hi = h[i], l = lw[i];
            nan_c += (hi != hi ? 1.f : 0.f) + (l != l ? 1.f : 0.f);
            inf_c += ((hi > 1e10f || hi < -1e10f) ? 1.f : 0.f) + (l > 1e10f ? 1.f : 0.f);
            if (finite_h(hi)) {
                hmin = fminf(hmin, hi);
                hmax = fmaxf(hmax, hi);
                sh += hi;
            }
            if (l == l && l < 1e10f) {
                const float w = expf(l - mx);
                sw += w;
                sw2 = fmaf(w, w, sw2);
            }
        }
        nan_c = block_reduce(nan_c, red, SumOp{});
        inf_c = block_reduce(inf_c, red, SumOp{});
        sw = block_reduce(sw, red, SumOp{});
        sw2 = block_reduce(sw2, red, SumOp{});
        hmin = block_reduce(hmin, red, MinOp{});
        hmax = block_reduce(hmax, red, MaxOp{});
        sh = block_reduce(sh, red, SumOp{});
        if (tid == 0) {
            out0[0] = nan_c;
            out0[1] = inf_c;
            out0[2] = sw;
            out0[3] = sw2;
            out0[4] = hmin;
            out0[5] = hmax;
            out0[6] = sh;
        }
        return;
    }
    // weighted means: the log-weight maximum over finite entries, then every
    // weight exp(lw - max) (a NaN log-weight makes the result NaN, as there)
    float mx = -1e30f;
    for (int i = tid; i < n; i += kDiagBS) {
        const float l = lw[i];
        if (!isnan(l) && !isinf(l)) mx = fmaxf(mx, l);
    }
    mx = block_reduce(mx, red, MaxOp{});
    float sg = 0.f, sw = 0.f, sq = 0.f;
    if (mode == SVPF_DIAG_NU) {
        // a = y_t, b = nu
        const float nu = b, inv_nu = 1.0f / nu, nu1 = nu + 1.0f;
        const float psi_d = svpf_digamma(0.5f * nu1) - svpf_digamma(0.5f * nu);
        for (int i = tid; i < n; i += kDiagBS) {
            const float w = expf(lw[i] - mx);
            const float z = a / (expf(0.5f * h[i]) + 1e-8f);
            const float z2 = z * z;
            const float A = fmaf(z2, inv_nu, 1.0f);
            const float g = 0.5f * (psi_d - inv_nu - logf(A) + nu1 * z2 / (nu * nu * A));
            sg = fmaf(w, g, sg);
            sw += w;
            sq = fmaf(w, z2, sq);
        }
    } else {
        // h = h_pred, h2 = h_prev; a = mu, b = rho, c = sigma
        const float inv_s = 1.0f / c, inv_s2 = inv_s * inv_s;
        for (int i = tid; i < n; i += kDiagBS) {
            const float w = expf(lw[i] - mx);
            const float eps = h[i] - a - b * (h2[i] - a);
            const float e2 = eps * eps * inv_s2;
            sg = fmaf(w, (e2 - 1.0f) * inv_s, sg);
            sw += w;
            sq = fmaf(w, e2, sq);
        }
    }
    sg = block_reduce(sg, red, SumOp{});
    sw = block_reduce(sw, red, SumOp{});
    sq = block_reduce(sq, red, SumOp{});
    if (tid == 0) {
        const float iw = 1.0f / (sw + 1e-8f);
        *out0 = sg * iw;
        *out1 = sq * iw;
    }
}

}  // namespace

extern "C" hipError_t svpf_launch_diag(int mode, const float* h, const float* h2, const float* lw, int n, float a, float b,
                                float c, float* out0, float* out1, hipStream_t stream) {
    if (n < 1 || !h || !lw || !out0 || (mode != SVPF_DIAG_HEALTH && !out1) || (mode == SVPF_DIAG_SIGMA && !h2) ||
        mode < SVPF_DIAG_HEALTH || mode > SVPF_DIAG_SIGMA)
        return hipErrorInvalidValue;
    svpf_diag_kernel<<<1, kDiagBS, 0, stream>>>(mode, h, h2, lw, n, a, b, c, out0, out1);
    return hipGetLastError();
}
