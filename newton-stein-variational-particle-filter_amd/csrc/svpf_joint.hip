// svpf_joint.hip -- gfx950 kernel of the joint state-parameter SVPF
// (svpf_joint/src/svpf_joint_kernels.cu, svpf_joint/src/svpf_joint.cu:196-376).
//
// One workgroup runs whole steps: predict with parameter diffusion, the four
// per-coordinate bandwidths, every Stein iteration and the diagnostics, with the
// particles held in registers across steps.  A Stein iteration writes one
// 32-byte record per particle to LDS -- the position scaled by 1/bandwidth per
// coordinate, then the score -- and every particle sums its pair terms over all
// records: the scaled squared distance is a sum of four squares, so a pair costs
// one reciprocal and eighteen FMA-class operations.
#include <hip/hip_runtime.h>

#include "svpf_device.hpp"
#include "svpf_kernels.h"

using namespace svpf;

namespace {

SVPF_DEV float jsigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }
SVPF_DEV float jclamp_h(float h) { return fmaxf(fminf(h, 5.0f), -15.0f); }
SVPF_DEV float jsafe_exp(float x) { return expf(fminf(fmaxf(x, -20.0f), 20.0f)); }

// The per-particle score and log weight (svpf_joint_kernels.cu:134-320).
SVPF_DEV void joint_grad(const svpf_jparams& P, float h, float hp, float mt, float rt, float st, float y_t, float g[4],
                         float& log_w) {
    const float rho = jsigmoid(rt);
    const float sigma = fmaxf(jsafe_exp(st), 0.01f);
    const float var = sigma * sigma;
    const float ysq = y_t * y_t;
    const float A = ysq / (jsafe_exp(h) + 1e-8f) / P.nu;
    log_w = P.st_const - 0.5f * h - (P.nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f));
    const float gh_lik = (logf(ysq + 1e-10f) - h + P.lik_offset) / 1.4f;
    // surprise of the observation against the particle's previous log-volatility
    const float z2o = ysq / (jsafe_exp(hp) + 1e-8f);
    const float inflate = 1.0f + 0.3f * fminf(fmaxf(z2o - 4.0f, 0.0f), 20.0f);
    float vm = fmaxf(var * inflate, 0.03f * 0.03f);
    const bool panic = z2o > 9.0f;
    if (panic) vm = fmaxf(vm, 1.0f);
    const float tw = panic ? 0.3f : 1.0f;
    const float diff = h - (mt + rho * (hp - mt));
    const float gh_tr = tw * (-diff / vm);
    const float gm_tr = tw * (diff / vm) * (1.0f - rho);
    const float gr_tr = tw * (diff / vm) * (hp - mt) * (rho * (1.0f - rho));
    const float gs_tr = (diff * diff) / var - 1.0f;
    float fobs = 0.0f;
    if (z2o > 9.0f)
        fobs = fminf((z2o - 9.0f) * 0.5f, 5.0f);
    else if (z2o < 1.5f)
        fobs = -1.0f * (1.5f - z2o);
    const float mr = -0.2f * (st - P.sig_pm);
    const float pm = -(mt - P.mu_pm) / P.mu_pv;
    const float pr = -(rt - P.rho_pm) / P.rho_pv;
    const float ps = -(st - P.sig_pm) / P.sig_pv;
    g[0] = fmaxf(fminf(gh_lik + gh_tr, 10.0f), -10.0f);
    g[1] = fmaxf(fminf(gm_tr + P.prior_weight * pm, 5.0f), -5.0f);
    g[2] = fmaxf(fminf(gr_tr + P.prior_weight * pr, 5.0f), -5.0f);
    g[3] = fmaxf(fminf(gs_tr + fobs + mr + P.prior_weight * ps, panic ? 10.0f : 5.0f), -5.0f);
}

// Workgroup sums of K <= 8 values for 16 waves.  Lane 0 of wave w stores its wave's
// sums at red[k * 16 + w]; then lane l of every wave loads red[l] and red[64 + l],
// a DPP scan inside each 16-lane row leaves the sum of value k in lane
// 16 * (k % 4) + 15 of the first (k < 4) or second (k >= 4) load.  Two barriers;
// every thread of the workgroup must call it.
template <int K>
SVPF_DEV void block_sum16(float (&v)[K], float* red) {
    static_assert(K >= 1 && K <= 8, "at most 8 values");
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[k * 16 + w] = v[k];
    }
    __syncthreads();
    float a = lane < K * 16 ? red[lane] : 0.0f;
    float b = 64 + lane < K * 16 ? red[64 + lane] : 0.0f;
    a += SVPF_DPP_F(0.0f, a, 0x111, 0xF);
    a += SVPF_DPP_F(0.0f, a, 0x112, 0xF);
    a += SVPF_DPP_F(0.0f, a, 0x114, 0xF);
    a += SVPF_DPP_F(0.0f, a, 0x118, 0xF);
    if constexpr (K > 4) {
        b += SVPF_DPP_F(0.0f, b, 0x111, 0xF);
        b += SVPF_DPP_F(0.0f, b, 0x112, 0xF);
        b += SVPF_DPP_F(0.0f, b, 0x114, 0xF);
        b += SVPF_DPP_F(0.0f, b, 0x118, 0xF);
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        v[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(k < 4 ? a : b), 16 * (k % 4) + 15));
    __syncthreads();
}

}  // namespace

// Thread t serves particles row + r * (BS / CS), r < R, with row = t / CS; the CS
// lanes of a row split the pair sum over i = col, col + CS, ...  The lanes of a row
// hold identical copies of its particles; lane col = 0 publishes and stores them.
//
// mode 0: T steps at ticks t0 .. t0 + T - 1 on observations y[0 .. T-1] (y1 when
// y is null and T = 1).
// mode 1: the initial draw (svpf_joint.cu:136-171) and its diagnostics.
//
// xs: [h | h_prev | mu~ | rho~ | sigma~], n floats each; gl: [g_h | g_mu | g_rho |
// g_sigma | log_w]; summ: param_mean[4], param_std[3], std_unconstrained[3],
// collapse_flags[3] (int), 3 unused words, then the last step's record
// (SVPF_JOINT_RECW words); rec: NULL or T records.
template <int BS, int R, int CS>
__global__ __launch_bounds__(BS) void svpf_joint_kernel(svpf_jparams P, float* __restrict__ xs,
                                                        float* __restrict__ gl, float* __restrict__ summ,
                                                        const float* __restrict__ y, float y1,
                                                        float* __restrict__ rec, int t0, int T, int mode) {
    constexpr int ROWS = BS / CS;
    constexpr int RC = R > 2 ? 2 : R;
    static_assert(R % RC == 0, "R must be a multiple of the chunk");
    static_assert(BS == 1024, "the reductions assume 16 waves");
    extern __shared__ __align__(16) char lds[];
    const int n = P.n;
    float4* pos = reinterpret_cast<float4*>(lds);
    float4* grd = pos + n;
    float* red = reinterpret_cast<float*>(grd + n);
    const int tid = threadIdx.x, row = tid / CS, col = tid % CS;
    const CounterRng rng{P.k0, P.k1, (uint32_t)P.filter};
    const float inv_n = 1.0f / (float)n;

    float h[R], hp[R], mt[R], rt[R], st[R], lw[R];
    bool valid[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = row + r * ROWS;
        valid[r] = j < n;
        lw[r] = 0.0f;
        if (mode == 1) {
            uint32_t a[4] = {0u, 0u, 0u, 0u}, b[4] = {0u, 0u, 0u, 0u};
            if (valid[r]) {
                rng.draw((uint32_t)j, 0u, SVPF_PURPOSE_JOINT_INIT_A, a);
                rng.draw((uint32_t)j, 0u, SVPF_PURPOSE_JOINT_INIT_B, b);
            }
            h[r] = P.mu_pm + 0.5f * normal_from_bits(a[0], a[1]);
            mt[r] = P.mu_pm + 0.2f * normal_from_bits(a[2], a[3]);
            rt[r] = P.rho_pm + 0.2f * normal_from_bits(b[0], b[1]);
            st[r] = P.sig_pm + 0.2f * normal_from_bits(b[2], b[3]);
            hp[r] = h[r];
        } else if (valid[r]) {
            h[r] = xs[j];
            hp[r] = xs[n + j];
            mt[r] = xs[2 * n + j];
            rt[r] = xs[3 * n + j];
            st[r] = xs[4 * n + j];
        } else {
            h[r] = hp[r] = mt[r] = rt[r] = st[r] = 0.0f;
        }
    }
    const bool owner = col == 0;
    float bw[4] = {0.0f, 0.0f, 0.0f, 0.0f};

    for (int k = 0; k < (mode == 1 ? 1 : T); ++k) {
        if (mode == 0) {
            const int t = t0 + k;
            const float y_t = y ? y[k] : y1;
            // ---- predict with parameter diffusion (svpf_joint_kernels.cu:79-128)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!valid[r]) continue;
                // an opaque copy of the index keeps the compiler from hoisting the
                // tick-invariant part of the generator's first round out of the tick
                // loop for every draw (it then spills them)
                uint32_t j = (uint32_t)(row + r * ROWS);
                __asm__ volatile("" : "+v"(j));
                uint32_t a[4], b[4];
                rng.draw(j, (uint32_t)t, SVPF_PURPOSE_JOINT_A, a);
                rng.draw(j, (uint32_t)t, SVPF_PURPOSE_JOINT_B, b);
                hp[r] = h[r];
                mt[r] = fmaxf(fminf(mt[r] + P.diff_mu * normal_from_bits(a[0], a[1]), 5.0f), -10.0f);
                rt[r] = fmaxf(fminf(rt[r] + P.diff_rho * normal_from_bits(a[2], a[3]), 5.0f), -5.0f);
                st[r] = fmaxf(fminf(st[r] + P.diff_sigma * normal_from_bits(b[0], b[1]), 2.0f), -5.0f);
                const float rho = jsigmoid(rt[r]);
                const float sigma = fmaxf(jsafe_exp(st[r]), 0.01f);
                const float h_pred = mt[r] + rho * (hp[r] - mt[r]);
                h[r] = jclamp_h(h_pred + sigma * normal_from_bits(b[2], b[3]));
            }
            // ---- per-coordinate bandwidths, sqrt(2 var / ln(n + 1)) clamped to [0.01, 2]
            // (svpf_joint_kernels.cu:620-663); the variance is two-pass.
            float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid[r] && owner) {
                    m[0] += h[r];
                    m[1] += mt[r];
                    m[2] += rt[r];
                    m[3] += st[r];
                }
            block_sum16<4>(m, red);
#pragma unroll
            for (int q = 0; q < 4; ++q) m[q] *= inv_n;
            float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid[r] && owner) {
                    const float d0 = h[r] - m[0], d1 = mt[r] - m[1], d2 = rt[r] - m[2], d3 = st[r] - m[3];
                    c[0] = fmaf(d0, d0, c[0]);
                    c[1] = fmaf(d1, d1, c[1]);
                    c[2] = fmaf(d2, d2, c[2]);
                    c[3] = fmaf(d3, d3, c[3]);
                }
            block_sum16<4>(c, red);
            float sc[4];
            const float inv_log = 1.0f / logf((float)n + 1.0f);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float bsq = fmaxf(2.0f * (c[q] * inv_n) * inv_log, 1e-6f);
                bw[q] = fmaxf(fminf(sqrtf(bsq), 2.0f), 0.01f);
                sc[q] = sqrtf(1.0f / (bw[q] * bw[q] + 1e-8f));
            }
            // surprise-boosted step sizes (svpf_joint_kernels.cu:384-406)
            bool sur[R];
            const float ysq = y_t * y_t;
#pragma unroll
            for (int r = 0; r < R; ++r) sur[r] = ysq / (expf(hp[r]) + 1e-8f) > 9.0f;
            const int n_iter = P.n_stein < 1 ? 0 : P.n_stein;
            if (n_iter == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (valid[r]) {
                        float g[4];
                        joint_grad(P, h[r], hp[r], mt[r], rt[r], st[r], y_t, g, lw[r]);
                        if (owner) grd[row + r * ROWS] = make_float4(g[0], g[1], g[2], g[3]);
                    }
            }
            // ---- Stein iterations: every particle reads the previous iteration's records
            for (int s = 0; s < n_iter; ++s) {
                float4 xj[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    xj[r] = make_float4(h[r] * sc[0], mt[r] * sc[1], rt[r] * sc[2], st[r] * sc[3]);
                    if (valid[r]) {
                        float g[4];
                        joint_grad(P, h[r], hp[r], mt[r], rt[r], st[r], y_t, g, lw[r]);
                        if (owner) {
                            const int j = row + r * ROWS;
                            pos[j] = xj[r];
                            grd[j] = make_float4(g[0], g[1], g[2], g[3]);
                        }
                    }
                }
                __syncthreads();
                // particles in chunks of RC: one pass over the records per chunk keeps
                // the accumulators of R = 4 within the register budget
#pragma unroll
                for (int c0 = 0; c0 < R; c0 += RC) {
                    float A[RC][4], B[RC][4], S[RC];
#pragma unroll
                    for (int u = 0; u < RC; ++u) {
                        S[u] = 0.0f;
#pragma unroll
                        for (int q = 0; q < 4; ++q) A[u][q] = B[u][q] = 0.0f;
                    }
                    for (int i = col; i < n; i += CS) {
                        const float4 p = pos[i];
                        const float4 gi = grd[i];
#pragma unroll
                        for (int u = 0; u < RC; ++u) {
                            const float4 xv = xj[c0 + u];
                            const float d0 = p.x - xv.x, d1 = p.y - xv.y, d2 = p.z - xv.z, d3 = p.w - xv.w;
                            const float base = fmaf(d0, d0, fmaf(d1, d1, fmaf(d2, d2, fmaf(d3, d3, 1.0f))));
                            const float K = __builtin_amdgcn_rcpf(base);
                            const float K2 = K * K;
                            A[u][0] = fmaf(K, gi.x, A[u][0]);
                            A[u][1] = fmaf(K, gi.y, A[u][1]);
                            A[u][2] = fmaf(K, gi.z, A[u][2]);
                            A[u][3] = fmaf(K, gi.w, A[u][3]);
                            B[u][0] = fmaf(K2, d0, B[u][0]);
                            B[u][1] = fmaf(K2, d1, B[u][1]);
                            B[u][2] = fmaf(K2, d2, B[u][2]);
                            B[u][3] = fmaf(K2, d3, B[u][3]);
                            S[u] += K;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < RC; ++u) {
                        const int r = c0 + u;
                        S[u] = group_sum<CS>(S[u]);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            A[u][q] = group_sum<CS>(A[u][q]);
                            B[u][q] = group_sum<CS>(B[u][q]);
                        }
                        // kernel-mass normalisation clamped to [32, n] (svpf_joint_kernels.cu:448-460);
                        // the repulsion sum_i -2 (x_i - x_j) / bw^2 K^2 in scaled coordinates.
                        const float inv_norm = 1.0f / fmaxf(fminf(S[u], (float)n), 32.0f);
                        float phi[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) phi[q] = (A[u][q] - 2.0f * sc[q] * B[u][q]) * inv_norm;
                        if (valid[r]) {
                            const float e_h = P.step_h * (sur[r] ? 5.0f : 1.0f);
                            const float e_mu = P.step_mu * (sur[r] ? 2.0f : 1.0f);
                            const float e_sigma = P.step_sigma * (sur[r] ? 4.0f : 1.0f);
                            h[r] = jclamp_h(h[r] + e_h * phi[0]);
                            mt[r] = fmaxf(fminf(mt[r] + e_mu * phi[1], 5.0f), -10.0f);
                            rt[r] = fmaxf(fminf(rt[r] + P.step_rho * phi[2], 5.0f), -5.0f);
                            st[r] = fmaxf(fminf(st[r] + e_sigma * phi[3], 2.0f), -5.0f);
                        }
                    }
                }
                __syncthreads();
            }
        }
        // ---- diagnostics (svpf_joint_kernels.cu:476-614, svpf_joint.cu:280-360)
        float s1[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float rv[R], sv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            rv[r] = jsigmoid(rt[r]);
            sv[r] = expf(st[r]);
            if (valid[r] && owner) {
                s1[0] += h[r];
                s1[1] += mt[r];
                s1[2] += rv[r];
                s1[3] += sv[r];
                s1[4] += rt[r];
                s1[5] += st[r];
            }
        }
        block_sum16<6>(s1, red);
        float mean[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) mean[q] = s1[q] * inv_n;
        float s2[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float mx = -3.402823466e38f;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid[r] && owner) {
                const float e1 = mt[r] - mean[1], e2 = rv[r] - mean[2], e3 = sv[r] - mean[3];
                const float e4 = rt[r] - mean[4], e5 = st[r] - mean[5];
                s2[0] = fmaf(e1, e1, s2[0]);
                s2[1] = fmaf(e2, e2, s2[1]);
                s2[2] = fmaf(e3, e3, s2[2]);
                s2[3] = fmaf(e4, e4, s2[3]);
                s2[4] = fmaf(e5, e5, s2[4]);
                mx = fmaxf(mx, lw[r]);
            }
        if (mode == 0) mx = block_max<BS>(mx, red);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid[r] && owner && mode == 0) {
                const float w = expf(lw[r] - mx);
                s2[5] += w;
                s2[6] = fmaf(w, w, s2[6]);
            }
        block_sum16<7>(s2, red);
        // Every value below is uniform across the workgroup: lane q of the first wave
        // writes record word q, lane 32 + q summary word q.
        if (tid < 64) {
            const float mu_std = sqrtf(fmaxf(s2[0] * inv_n, 0.0f));
            const float std_rt = sqrtf(fmaxf(s2[3] * inv_n, 0.0f)), std_st = sqrtf(fmaxf(s2[4] * inv_n, 0.0f));
            const int lane = tid & 31;
            // summary words: param_mean (mu, rho, sigma, h), param_std, std_unconstrained, flags
            constexpr int kSumm[13] = {2, 3, 4, 0, 5, 6, 7, 8, 9, 10, 11, 12, 13};
            int q = lane;
            if (tid >= 32) q = lane < 13 ? kSumm[lane] : -1;
            float v = 0.0f;
            switch (q) {
                case 0: v = mean[0]; break;
                case 1: v = expf(mean[0] * 0.5f); break;
                case 2: v = mean[1]; break;
                case 3: v = mean[2]; break;
                case 4: v = mean[3]; break;
                case 5: v = mu_std; break;
                case 6: v = sqrtf(fmaxf(s2[1] * inv_n, 0.0f)); break;
                case 7: v = sqrtf(fmaxf(s2[2] * inv_n, 0.0f)); break;
                case 8: v = mu_std; break;
                case 9: v = std_rt; break;
                case 10: v = std_st; break;
                case 11: v = __int_as_float(mu_std < P.thr_mu ? 1 : 0); break;
                case 12: v = __int_as_float(std_rt < P.thr_rho ? 1 : 0); break;
                case 13: v = __int_as_float(std_st < P.thr_sigma ? 1 : 0); break;
                case 14: v = mode == 0 ? (s2[5] * s2[5]) / (s2[6] + 1e-10f) : 0.0f; break;
                case 15: v = mode == 0 ? mx + logf(s2[5]) - logf((float)n) : 0.0f; break;
                case 16: v = bw[0]; break;
                case 17: v = bw[1]; break;
                case 18: v = bw[2]; break;
                case 19: v = bw[3]; break;
                default: break;
            }
            const bool last = k == (mode == 1 ? 0 : T - 1);
            if (tid < 32) {
                if (lane < SVPF_JOINT_RECW_DEV) {
                    if (rec) rec[(size_t)k * SVPF_JOINT_RECW_DEV + lane] = v;
                    if (last) summ[16 + lane] = v;
                }
            } else if (last && q >= 0) {
                summ[lane] = v;
            }
        }
    }
    // ---- store the particles and the last iteration's scores (still in the LDS
    // records) and log weights
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = row + r * ROWS;
        if (!valid[r] || !owner) continue;
        xs[j] = h[r];
        xs[n + j] = hp[r];
        xs[2 * n + j] = mt[r];
        xs[3 * n + j] = rt[r];
        xs[4 * n + j] = st[r];
        const float4 g = mode == 0 ? grd[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        gl[j] = g.x;
        gl[n + j] = g.y;
        gl[2 * n + j] = g.z;
        gl[3 * n + j] = g.w;
        gl[4 * n + j] = lw[r];
    }
}

namespace {
struct JointTile {
    const void* fn;
    int bs;
};

JointTile joint_tile(int n) {
#define SVPF_KJ(BSV, RV, CSV) return JointTile{(const void*)svpf_joint_kernel<BSV, RV, CSV>, BSV}
    if (n <= 128) SVPF_KJ(1024, 1, 8);
    if (n <= 256) SVPF_KJ(1024, 1, 4);
    if (n <= 512) SVPF_KJ(1024, 1, 2);
    if (n <= 1024) SVPF_KJ(1024, 1, 1);
    if (n <= 2048) SVPF_KJ(1024, 2, 1);
    if (n <= 4096) SVPF_KJ(1024, 4, 1);
    return JointTile{nullptr, 0};
#undef SVPF_KJ
}
}  // namespace

extern "C" size_t svpf_joint_lds_bytes_host(int n) {
    // two float4 records per particle and the reduction scratch (8 values x 16 waves)
    return (size_t)n * 32 + 8 * 16 * sizeof(float);
}

extern "C" hipError_t svpf_launch_joint(const svpf_jparams* P, float* xs, float* gl, float* summ, const float* y,
                                        float y1, float* rec, int t0, int T, int mode, hipStream_t stream) {
    const JointTile tl = joint_tile(P->n);
    if (!tl.fn || P->n < 1 || (mode == 0 && (T < 1 || (!y && T != 1)))) return hipErrorInvalidValue;
    const size_t lds = svpf_joint_lds_bytes_host(P->n);
    hipError_t e = hipFuncSetAttribute(tl.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    svpf_jparams pv = *P;
    void* args[] = {(void*)&pv,   (void*)&xs,  (void*)&gl, (void*)&summ, (void*)&y,
                    (void*)&y1,   (void*)&rec, (void*)&t0, (void*)&T,    (void*)&mode};
    e = hipLaunchKernel(tl.fn, dim3(1), dim3(tl.bs), args, lds, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}
