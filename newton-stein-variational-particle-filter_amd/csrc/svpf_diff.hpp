// svpf_diff.hpp -- the differentiable SVPF: a forward that keeps checkpoints
// and a hand-written adjoint.
//
// Model (one filter, per tick t, positions x from the previous tick, RMSProp
// state v; src/svpf_opt_kernels.cu:125-155, :354-521, :742-906,
// src/svpf_optimized_graph.cu:1057-1317):
//   m_i   = mu + rho (x_i - mu) + gamma y_{t-1} / (e^{x_i/2} + 1e-8)
//   h_i   = clamp(m_i + sigma z_i)                              predict
//   ll_t  = logmeanexp_i log p(y_t | h_i)                       predictive loglik (the loss terms)
//   ell^2 = lower median of (h_i - h_j)^2, bw = clamp(ell, .01, 2)
//   S times: g_j = clamp(g_prior_j + beta_k g_lik_j), curv_j = clamp(-H_lik - H_prior),
//            s_j = 0.7 g_j / curv_j, c_j = +-2 / curv_j,
//            phi_i = (1/N) sum_j K_ij s_j + c_j (h_i - h_j)/bw^2 K_ij^2,  K = 1/(1 + (h_i-h_j)^2/bw^2)
//            v_i <- rho_v v_i + (1-rho_v) phi_i^2,  h_i <- clamp(h_i + eps_k phi_i / sqrt(v_i + e) + noise)
//   with the mixture prior g_prior_j = -sum_i r_ij (h_j - m_i) / sigma^2 / sum_i r_ij,
//   r_ij = exp(-(h_j - m_i)^2 / 2 sigma^2).
//
// Forward: one workgroup per filter runs all T ticks with the particles in
// registers and LDS, and stores (x, v) every C ticks.
//
// Backward: per checkpoint segment (last first) the ticks are recomputed from
// the checkpoint by the same forward code, which writes a per-tick tape to
// HBM (positions, RMSProp state, phi and the responsibility moments of every
// iteration); then the ticks are reversed.  Per Stein iteration the adjoint
// costs two O(N^2) passes:
//   pass C (transport): one sweep over particle pairs accumulates the row sums
//     sum_j s_j d K^2 and sum_j c_j K^2 (1 - 4 d^2 K) and, with the adjoint
//     weights w_j, the column sums sum_j w_j K, sum_j w_j d K^2 and
//     sum_j w_j K^2 (1 - 4 d^2 K) (K symmetric, d antisymmetric, so a column
//     sum is a row sum of w); the bandwidth adjoint follows from these by
//     translation invariance, with no further pair sums;
//   pass B (mixture prior): every component i sums over particles j
//     a_j rt_ij (1 - d_ij (d_ij - E_j d) / sigma^2), rt the normalised
//     responsibilities rebuilt from the taped log-normaliser.
// Positions are scaled as in the filter engine: x/bw in pass C, and
// sqrt(log2 e / 2 sigma^2) (h - m) in the prior, where a responsibility is one exp2.
#pragma once

#include "svpf_device.hpp"
#include "svpf_special.h"

namespace svpf {

constexpr float kDiffLog2e = 1.4426950408889634f;
constexpr float kDiffKappa = 1.3862943611198906f;  // 2 ln 2 = sigma^-2 / (scale^2) ratio of the prior units
constexpr float kDiffAlpha = 0.8493218002880191f;  // sqrt(log2 e / 2)

// Tape slot of one tick, in floats: x [n], u (predicted, before the clamp) [n],
// then per iteration the TP_ROWS rows of n floats, then 4 scalars.
// (The prior's third moment is not taped: its sigma adjoint follows from homogeneity, see tick_back.)
enum { TP_H = 0, TP_V = 1, TP_PHI = 2, TP_EDD = 3, TP_VDD = 4, TP_L = 5, TP_UN = 6, TP_ROWS = 7 };
__host__ __device__ inline size_t diff_slot_floats(int n, int S) { return (size_t)(2 + TP_ROWS * S) * n + 4; }

struct DiffTheta {
    float mu, rho, sigma, nu;
    float st_const;   // log normaliser of the Student-t density (src/svpf_optimized_graph.cu:123-125)
    float dst_const;  // its nu-derivative
};

SVPF_DEV DiffTheta diff_theta(const float* th) {
    DiffTheta t;
    t.mu = th[0];
    t.rho = th[1];
    t.sigma = th[2];
    t.nu = th[3];
    t.st_const = lgammaf((t.nu + 1.0f) * 0.5f) - lgammaf(t.nu * 0.5f) - 0.5f * logf(3.14159265358979f * t.nu);
    t.dst_const = 0.5f * svpf_digamma((t.nu + 1.0f) * 0.5f) - 0.5f * svpf_digamma(t.nu * 0.5f) - 0.5f / t.nu;
    return t;
}

struct DiffLds {
    __host__ __device__ static int mbuf_len(int n) {
        int p = 4;
        while (p < n + 1) p <<= 1;
        return p;
    }
    // the record area also holds the final 4 x 16 doubles of the theta-adjoint reduction
    __host__ __device__ static int rec_len(int nmax) { return nmax < 32 ? 32 : nmax; }
    __host__ __device__ static size_t bytes(int nmax) {
        size_t b = (size_t)(rec_len(nmax) + nmax) * 16 + (size_t)mbuf_len(nmax) * 4 + kRedWords * 4 + 16 * 4;
        return (b + 15) & ~(size_t)15;
    }
};

template <int BS, int R>
struct DiffEngine {
    float4* rec;    // [nmax] {h/bw, s, c/bw, w}; sort input and median scratch in the forward bandwidth
    float4* recb;   // [nmax] pass-B records {scaled h, log2 normaliser, a, kappa E dd}; predicted h in the pair
                    // search; the sorted means during the forward Stein loop
    float* mbuf;    // [p2] scaled mixture means (padded with far sentinels) / sorted positions
    float* red;     // [kRedWords] reduction scratch
    float* scal;    // [16] broadcast scalars
    int n, p2;
    float inv_n;
    float x[R], v[R];  // positions and RMSProp state between ticks

    SVPF_DEV void bind(char* lds, int n_, int nmax) {
        n = __builtin_amdgcn_readfirstlane(n_);
        p2 = DiffLds::mbuf_len(n);
        rec = reinterpret_cast<float4*>(lds);
        recb = rec + DiffLds::rec_len(nmax);
        mbuf = reinterpret_cast<float*>(recb + nmax);
        red = mbuf + DiffLds::mbuf_len(nmax);
        scal = red + kRedWords;
        inv_n = 1.0f / (float)n;
    }
    SVPF_DEV static int row(int r) { return (int)threadIdx.x + r * BS; }
    SVPF_DEV bool valid(int r) const { return row(r) < n; }

    SVPF_DEV float noise_z(const svpf_ddesc& D, int t, int i) const {
        if (D.z) return D.z[(size_t)t * n + i];
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        uint32_t o[4];
        rng.draw((uint32_t)i, (uint32_t)t, SVPF_PURPOSE_PREDICT, o);
        return normal_from_bits(o[0], o[1]);
    }
    SVPF_DEV float noise_xi(const svpf_ddesc& D, int S, int t, int k, int i) const {
        if (D.z) return D.xi[((size_t)t * S + k) * n + i];
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        return rng.normal((uint32_t)i, (uint32_t)t, SVPF_PURPOSE_SVLD0 + (uint32_t)k);
    }

    SVPF_DEV float logmeanexp(const float (&a)[R]) {
        float mx = -3.0e38f;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) mx = fmaxf(mx, a[r]);
        mx = block_max<BS>(mx, red);
        float se[1] = {0.f};
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) se[0] += expf(a[r] - mx);
        block_sum<BS, 1>(se, red);
        return mx + logf(fmaxf(se[0] * inv_n, 1e-10f));
    }

    // ---- sorted copy of the means (recb is free during the forward Stein loop) ----
    // The means do not change during a tick's Stein loop.  With a sorted copy
    // the shift min_i fl(hh - mm_i)^2 (the reference's max term) is a binary
    // search: rounding is monotone, so it is attained at one of the two means
    // around hh and is bit-identical to the minimum over all of them.  The
    // sums still run over mbuf in its own order.
    SVPF_DEV void sort_means_copy() {
        float* src = reinterpret_cast<float*>(rec);
        __syncthreads();  // mbuf complete
        for (int i = (int)threadIdx.x; i < n; i += BS) src[i] = mbuf[i];
        __syncthreads();
        block_sort<BS>(src, reinterpret_cast<float*>(recb), n);
    }
    SVPF_DEV float nearest_sq(float hh) const {
        const float* ms = reinterpret_cast<const float*>(recb);
        int pos = -1;  // last index with ms[pos] < hh
        for (int step = 1 << (31 - __clz(n)); step > 0; step >>= 1) {
            const int q = pos + step;
            pos = (q < n && ms[q] < hh) ? q : pos;
        }
        float best = 3.0e38f;
        if (pos >= 0) {
            const float d = hh - ms[pos];
            best = d * d;
        }
        if (pos + 1 < n) {
            const float d = hh - ms[pos + 1];
            best = fminf(best, d * d);
        }
        return best;
    }

    // ---- mixture-prior sums of own rows over all components (scaled units) ----
    template <bool MOM>
    SVPF_DEV void prior_sums(const float (&hh)[R], float (&off)[R], float (&sr)[R], float (&s1)[R],
                             float (&s2)[R]) const {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            off[r] = nearest_sq(hh[r]);
            sr[r] = s1[r] = s2[r] = 0.f;
        }
        const int nq = (n + 3) >> 2;
        const float4* m4 = reinterpret_cast<const float4*>(mbuf);
        constexpr int U = 2;
        int q = 0;
        auto acc = [&](const float4& m) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dd[4] = {hh[r] - m.x, hh[r] - m.y, hh[r] - m.z, hh[r] - m.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float e = __builtin_amdgcn_exp2f(fmaf(-dd[c], dd[c], off[r]));
                    const float ed = e * dd[c];
                    sr[r] += e;
                    s1[r] += ed;
                    if (MOM) s2[r] = fmaf(ed, dd[c], s2[r]);
                }
            }
        };
        q = 0;
        for (; q + U <= nq; q += U) {
            float4 m[U];
#pragma unroll
            for (int u = 0; u < U; ++u) m[u] = m4[q + u];
#pragma unroll
            for (int u = 0; u < U; ++u) acc(m[u]);
        }
        for (; q < nq; ++q) acc(m4[q]);
    }

    // ---- IMQ transport of own rows: acc_i = sum_j K (s_j + cr_j d K) ----
    SVPF_DEV void transport(const float (&xs)[R], float (&acc)[R]) const {
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = 0.f;
        constexpr int U = R >= 2 ? 2 : 4;
        auto body = [&](const float4& e) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float d = xs[r] - e.x;
                const float K = __builtin_amdgcn_rcpf(fmaf(d, d, 1.0f));
                acc[r] = fmaf(K, fmaf(e.z * d, K, e.y), acc[r]);
            }
        };
        int j = 0;
        for (; j + U <= n; j += U) {
            float4 e[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = rec[j + u];
#pragma unroll
            for (int u = 0; u < U; ++u) body(e[u]);
        }
        for (; j < n; ++j) body(rec[j]);
    }

    // ---- bandwidth: lower median of pairwise r^2 of own rows' h (sort + selection) ----
    SVPF_DEV float median_ell2(const float (&h)[R], float est) {
        float* tmp = reinterpret_cast<float*>(rec);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) tmp[row(r)] = h[r];
        __syncthreads();
        block_sort<BS>(tmp, mbuf, n);
        int* plo = reinterpret_cast<int*>(rec);
        constexpr int KR = R < 2 ? 2 : R;
        float med = median_pairwise_r2<BS, KR>(mbuf, n, est, plo, plo + n, reinterpret_cast<float*>(plo + 2 * n),
                                               2 * n, reinterpret_cast<uint32_t*>(red));
        __syncthreads();
        return med;
    }

    // ---- the pair (a, b), a < b, whose squared distance is ell2 (first in row-major order).
    //      sorted: mbuf holds the values ascending (the median's sort); otherwise they are sorted here ----
    SVPF_DEV void median_pair(const float (&h)[R], float ell2, bool sorted, int& a, int& b) {
        if (!sorted) {
            float* tmp = reinterpret_cast<float*>(rec);
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) tmp[row(r)] = h[r];
            __syncthreads();
            block_sort<BS>(tmp, mbuf, n);
        }
        float* hl = reinterpret_cast<float*>(recb);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) hl[row(r)] = h[r];
        first_pair_at_r2<BS, R>(h, hl, mbuf, n, ell2, reinterpret_cast<int*>(hl + n), reinterpret_cast<int*>(scal + 8),
                                a, b);
    }

    // ---- one tick forward.  SELECT: select ell2 (the forward) or read the
    //      forward's value (the recompute).  RECORD: write the tape slot. ----
    template <bool RECORD, bool SELECT>
    SVPF_DEV float tick(const svpf_dconfig& C, const DiffTheta& th, const svpf_ddesc& D, int t, float y_prev,
                        float* slot, float& med_est) {
        const int S = C.S;
        const float y_t = D.y[t];
        const float mu = th.mu, rho = th.rho, sigma = th.sigma, nu = th.nu;
        float m[R], h[R], lw[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            const bool ok = i < n;
            const float z = ok ? noise_z(D, t, i) : 0.f;
            const float xi_ = x[r];
            const float lev = C.gamma * y_prev / (safe_expf(xi_ * 0.5f) + 1e-8f);
            // explicit FMAs here and in the update below: the forward and the checkpoint recompute are
            // two instantiations of this code, and contraction left to the compiler may round them apart
            // (the recompute must reproduce the forward's positions bit for bit)
            m[r] = fmaf(rho, xi_ - mu, mu) + lev;
            const float u = fmaf(sigma, z, m[r]);
            h[r] = clamp_logvol(u);
            if (RECORD && ok) {
                slot[i] = xi_;
                slot[n + i] = u;
            }
            const float A = (y_t * y_t) / (safe_expf(h[r]) + 1e-8f) / nu;
            lw[r] = ok ? th.st_const - 0.5f * h[r] - (nu + 1.0f) * 0.5f * log1pf(A) : -3.0e38f;
        }
        const float pred_ll = logmeanexp(lw);

        float ell2;
        if (SELECT) {
            ell2 = median_ell2(h, med_est);
            med_est = ell2;
            if (threadIdx.x == 0) D.ell2[t] = ell2;
        } else {
            ell2 = D.ell2[t];
        }
        if (RECORD) {
            int a, b;
            const float bwr = sqrtf(ell2);
            if (bwr > 0.01f && bwr < 2.0f) {
                median_pair(h, ell2, SELECT, a, b);
            } else {
                a = b = -1;
            }
            if (threadIdx.x == 0) {
                float* sc = slot + (size_t)(2 + TP_ROWS * S) * n;
                sc[0] = ell2;
                sc[1] = __int_as_float(a);
                sc[2] = __int_as_float(b);
                if (D.mpair) {
                    D.mpair[2 * t] = a;
                    D.mpair[2 * t + 1] = b;
                }
            }
        }
        const float bw = fminf(fmaxf(sqrtf(ell2), 0.01f), 2.0f);
        const float rs = 1.0f / bw;

        const float inv_s2 = 1.0f / (sigma * sigma);
        const float sq = sqrtf(kDiffLog2e * 0.5f * inv_s2);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) mbuf[row(r)] = sq * m[r];
        for (int i = n + (int)threadIdx.x; i < p2; i += BS) mbuf[i] = 1.0e18f;  // exp2 underflows to 0
        sort_means_copy();

        const float sgn2 = C.paper_sign ? 2.0f : -2.0f;
        const float temp = C.temperature;
        const float ysq = y_t * y_t;
        const float isq = inv_s2 / sq;
        float* it = slot + 2 * (size_t)n;
        // the recompute with a stored drift replays the positions: no records, no transport pass
        const bool replay = RECORD && !SELECT && D.drift != nullptr;
        for (int k = 0; k < S; ++k, it += TP_ROWS * (size_t)n) {
            const float beta = S > 1 ? 0.3f + 0.7f * (float)k / (float)(S - 1) : 1.0f;
            const float eff = 0.1f * sqrtf(beta);
            const float dstd = sqrtf(2.0f * eff * temp);
            __syncthreads();  // means complete; the previous transport finished reading rec
            float hh[R], off[R], sr[R], s1[R], s2[R];
#pragma unroll
            for (int r = 0; r < R; ++r) hh[r] = h[r] * sq;
            prior_sums<RECORD>(hh, off, sr, s1, s2);
            float xs[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                xs[r] = h[r] * rs;
                if (!replay) {
                    const float gp = -((s1[r] * isq) / (sr[r] + 1e-8f));
                    const float A = ysq / (safe_expf(h[r]) + 1e-8f) / nu;
                    const float opa = 1.0f + A;
                    const float gl = -0.5f + 0.5f * (nu + 1.0f) * A / opa - C.lik_offset;
                    const float g = fminf(fmaxf(gp + beta * gl, -10.0f), 10.0f);
                    const float cv = fminf(fmaxf(0.5f * (nu + 1.0f) * A / (opa * opa) + inv_s2, 0.1f), 100.0f);
                    const float s = 0.7f * g / cv;
                    if (i < n) rec[i] = make_float4(xs[r], s, (sgn2 / cv) * rs, 0.0f);
                }
                if (i < n) {
                    if (RECORD) {
                        const float Edd = s1[r] / sr[r];
                        const float m2 = s2[r] / sr[r];
                        it[TP_H * (size_t)n + i] = h[r];
                        it[TP_V * (size_t)n + i] = v[r];
                        it[TP_EDD * (size_t)n + i] = Edd;
                        it[TP_VDD * (size_t)n + i] = m2 - Edd * Edd;
                        it[TP_L * (size_t)n + i] = __log2f(sr[r]) - off[r];
                    }
                }
            }
            float acc[R];
            if (!replay) {
                __syncthreads();  // records complete
                transport(xs, acc);
            }
            float* drift_k = D.drift ? D.drift + ((size_t)t * S + k) * n : nullptr;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                const float p = replay ? (i < n ? drift_k[i] : 0.0f) : acc[r] * inv_n;
                if (SELECT && drift_k && i < n) drift_k[i] = p;
                const float vn = fmaf(C.rms_rho, v[r], (1.0f - C.rms_rho) * (p * p));
                const float xi = (temp > 1e-6f && i < n) ? noise_xi(D, S, t, k, i) : 0.0f;
                const float un = fmaf(dstd, xi, fmaf(eff * p, rsqrtf(vn + C.rms_eps), h[r]));
                if (RECORD && i < n) {
                    it[TP_PHI * (size_t)n + i] = p;
                    it[TP_UN * (size_t)n + i] = un;
                }
                v[r] = vn;
                h[r] = clamp_logvol(un);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = h[r];
        return pred_ll;
    }

    // ---- pass C: adjoint of the transport over all pairs ----
    SVPF_DEV void pass_c(const float (&xs)[R], float (&R1)[R], float (&R2)[R], float (&C1)[R], float (&C2)[R],
                         float (&C3)[R]) const {
#pragma unroll
        for (int r = 0; r < R; ++r) R1[r] = R2[r] = C1[r] = C2[r] = C3[r] = 0.f;
        constexpr int U = R >= 2 ? 1 : 2;
        auto body = [&](const float4& e) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float d = xs[r] - e.x;
                const float K = __builtin_amdgcn_rcpf(fmaf(d, d, 1.0f));
                const float K2 = K * K;
                const float dK2 = d * K2;
                const float tt = K2 * fmaf(4.0f, K, -3.0f);  // K^2 (1 - 4 d^2 K), d^2 K = 1 - K
                R1[r] = fmaf(e.y, dK2, R1[r]);
                R2[r] = fmaf(e.z, tt, R2[r]);
                C1[r] = fmaf(e.w, K, C1[r]);
                C2[r] = fmaf(e.w, dK2, C2[r]);
                C3[r] = fmaf(e.w, tt, C3[r]);
            }
        };
        int j = 0;
        for (; j + U <= n; j += U) {
            float4 e[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = rec[j + u];
#pragma unroll
            for (int u = 0; u < U; ++u) body(e[u]);
        }
        for (; j < n; ++j) body(rec[j]);
    }

    // ---- pass B: component adjoints mb_i += sum_j a_j rt_ij (1 - kappa dd (dd - Edd_j)) ----
    SVPF_DEV void pass_b(const float (&mm)[R], float (&mb)[R]) const {
        float acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = 0.f;
        constexpr int U = R >= 2 ? 1 : 2;
        auto body = [&](const float4& e) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dd = e.x - mm[r];
                const float rt = __builtin_amdgcn_exp2f(fmaf(-dd, dd, -e.y));
                const float ar = e.z * rt;
                // 1 - kappa dd (dd - Edd) = 1 + dd (kappa Edd - kappa dd); e.w = kappa Edd
                acc[r] = fmaf(ar, fmaf(fmaf(-kDiffKappa, dd, e.w), dd, 1.0f), acc[r]);
            }
        };
        int j = 0;
        for (; j + U <= n; j += U) {
            float4 e[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = recb[j + u];
#pragma unroll
            for (int u = 0; u < U; ++u) body(e[u]);
        }
        for (; j < n; ++j) body(recb[j]);
#pragma unroll
        for (int r = 0; r < R; ++r) mb[r] += acc[r];
    }

    // ---- one tick backward.  xb, vb: adjoints of (x, v) after the tick, replaced
    //      by those before it; tb accumulates dL/dtheta of own rows. ----
    SVPF_DEV void tick_back(const svpf_dconfig& C, const DiffTheta& th, const svpf_ddesc& D, int t, float y_prev,
                            const float* slot, float glbar, float (&xb)[R], float (&vb)[R], double (&tb)[4]) {
        const int S = C.S;
        const float y_t = D.y[t];
        const float mu = th.mu, rho = th.rho, sigma = th.sigma, nu = th.nu;
        const float* sc = slot + (size_t)(2 + TP_ROWS * S) * n;
        if (threadIdx.x == 0) {
            scal[0] = sc[0];
            scal[1] = sc[1];
            scal[2] = sc[2];
        }
        __syncthreads();
        const float ell2 = scal[0];
        const int pa = __float_as_int(scal[1]), pb = __float_as_int(scal[2]);
        const float bw = fminf(fmaxf(sqrtf(ell2), 0.01f), 2.0f);
        const float rs = 1.0f / bw;
        const float inv_s2 = 1.0f / (sigma * sigma);
        const float sq = sqrtf(kDiffLog2e * 0.5f * inv_s2);
        const float isq = inv_s2 / sq;
        const float isq_m = 1.0f / sq, isig = 1.0f / sigma;
        const float sgn2 = C.paper_sign ? 2.0f : -2.0f;
        const float ysq = y_t * y_t;
        const float np1 = nu + 1.0f;

        float xr[R], u[R], mm[R], mb[R], hb[R], vbar[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            const bool ok = i < n;
            xr[r] = ok ? slot[i] : 0.f;
            u[r] = ok ? slot[n + i] : 0.f;
            const float lev = C.gamma * y_prev / (safe_expf(xr[r] * 0.5f) + 1e-8f);
            mm[r] = sq * (fmaf(rho, xr[r] - mu, mu) + lev);  // the forward's means, rounded alike
            mb[r] = 0.f;
            hb[r] = ok ? xb[r] : 0.f;
            vbar[r] = ok ? vb[r] : 0.f;
        }
        float rsb = 0.f;
        double ts = 0.0, tn = 0.0;  // sigma and nu adjoints of own rows
        for (int k = S - 1; k >= 0; --k) {
            const float* it = slot + 2 * (size_t)n + (size_t)k * TP_ROWS * n;
            const float beta = S > 1 ? 0.3f + 0.7f * (float)k / (float)(S - 1) : 1.0f;
            const float eff = 0.1f * sqrtf(beta);
            float h[R], s[R], c[R], w[R], xs[R], curv[R], gmask[R], cmask[R], A[R], E[R], Edd[R], Vdd[R], L[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                const bool ok = i < n;
                h[r] = ok ? it[TP_H * (size_t)n + i] : 0.f;
                const float vv = ok ? it[TP_V * (size_t)n + i] : 0.f;
                const float p = ok ? it[TP_PHI * (size_t)n + i] : 0.f;
                Edd[r] = ok ? it[TP_EDD * (size_t)n + i] : 0.f;
                Vdd[r] = ok ? it[TP_VDD * (size_t)n + i] : 0.f;
                L[r] = ok ? it[TP_L * (size_t)n + i] : 0.f;
                const float un = ok ? it[TP_UN * (size_t)n + i] : 0.f;
                E[r] = safe_expf(h[r]);
                A[r] = ysq / (E[r] + 1e-8f) / nu;
                const float opa = 1.0f + A[r];
                const float gp = -Edd[r] * isq;
                const float gl = -0.5f + 0.5f * np1 * A[r] / opa - C.lik_offset;
                const float graw = gp + beta * gl;
                gmask[r] = (graw > -10.0f && graw < 10.0f) ? 1.0f : 0.0f;
                const float g = fminf(fmaxf(graw, -10.0f), 10.0f);
                const float craw = 0.5f * np1 * A[r] / (opa * opa) + inv_s2;
                cmask[r] = (craw > 0.1f && craw < 100.0f) ? 1.0f : 0.0f;
                curv[r] = fminf(fmaxf(craw, 0.1f), 100.0f);
                s[r] = 0.7f * g / curv[r];
                c[r] = sgn2 / curv[r];
                // RMSProp + clamp
                const float vn = fmaf(C.rms_rho, vv, (1.0f - C.rms_rho) * (p * p));
                const float q = rsqrtf(vn + C.rms_eps);
                const float hm = (ok && un > -15.0f && un < 5.0f) ? 1.0f : 0.0f;
                const float hbb = hb[r] * hm;
                const float vt = vbar[r] + hbb * (-0.5f * eff * p * q * q * q);
                w[r] = ok ? (hbb * eff * q + vt * 2.0f * (1.0f - C.rms_rho) * p) * inv_n : 0.f;
                vbar[r] = C.rms_rho * vt;
                hb[r] = hbb;
                xs[r] = h[r] * rs;
            }
            __syncthreads();  // the previous pass C / pass B finished reading rec / recb
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) rec[row(r)] = make_float4(xs[r], s[r], c[r] * rs, w[r]);
            __syncthreads();
            float R1[R], R2[R], C1[R], C2[R], C3[R];
            pass_c(xs, R1, R2, C1, C2, C3);
            float a[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                a[r] = 0.f;
                if (!valid(r)) continue;  // rows past n see nonzero column sums: keep them out of the adjoints
                const float cr = c[r] * rs;
                // G = d(sum_i w_i phi_i)/d(h_i / bw); the transport sees positions only through
                // differences, so the bandwidth adjoint is bw (sum_i x_i G_i - sum_i cr_i C2_i)
                const float G = w[r] * (R2[r] - 2.0f * R1[r]) - 2.0f * s[r] * C2[r] - cr * C3[r];
                hb[r] += rs * G;
                rsb += (xs[r] * G - cr * C2[r]) * bw;
                const float sbar = C1[r];
                const float cbar = -C2[r] * rs;
                const float gbar = gmask[r] * sbar * 0.7f / curv[r];
                const float cvbar = -cmask[r] * (sbar * s[r] + cbar * c[r]) / curv[r];
                const float opa = 1.0f + A[r];
                const float glb = beta * gbar;
                const float Abar = glb * 0.5f * np1 / (opa * opa) + cvbar * 0.5f * np1 * (1.0f - A[r]) / (opa * opa * opa);
                tn += (double)(glb * 0.5f * A[r] / opa + cvbar * 0.5f * A[r] / (opa * opa) - Abar * A[r] / nu);
                ts += (double)(cvbar * (-2.0f * inv_s2 / sigma));
                if (h[r] < 20.0f) hb[r] += Abar * (-A[r] * E[r] / (E[r] + 1e-8f));
                const float hp_ = gbar * inv_s2 * (kDiffKappa * Vdd[r] - 1.0f);
                hb[r] += hp_;
                // d gp / d sigma by homogeneity: scaling every particle, every mean and sigma by lambda scales
                // gp by 1/lambda, so sigma d(gbar gp)/dsigma = -gbar gp - h hbar - sum_j m_j mbar_j with hbar,
                // mbar the prior's position and mean adjoints (pass B); the means' part follows the Stein loop
                ts -= (double)((gbar * (-Edd[r] * isq) + h[r] * hp_) * isig);
                a[r] = gbar * inv_s2;
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) recb[row(r)] = make_float4(h[r] * sq, L[r], a[r], kDiffKappa * Edd[r]);
            __syncthreads();
            pass_b(mm, mb);
        }
        // bandwidth: ell^2 = (h_a - h_b)^2 of the predicted particles
        float hp[R];
#pragma unroll
        for (int r = 0; r < R; ++r) hp[r] = clamp_logvol(u[r]);
        float rr[1] = {rsb};
        block_sum<BS, 1>(rr, red);  // also orders the last pass B before recb is reused
        const float sb = sqrtf(ell2);
        if (pa >= 0 && sb > 0.01f && sb < 2.0f) {
            float* hl = reinterpret_cast<float*>(recb);
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) hl[row(r)] = hp[r];
            __syncthreads();
            // rsb is the adjoint of rs = ell^-1, and d rs / d ell^2 = -rs^3 / 2
            const float l2bar = rr[0] * (-0.5f * rs * rs * rs);
            const float dab = hl[pa] - hl[pb];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (row(r) == pa) hb[r] += 2.0f * dab * l2bar;
                if (row(r) == pb) hb[r] -= 2.0f * dab * l2bar;
            }
        }
        // predictive log-likelihood
        float lw[R], Ep[R], Ap[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Ep[r] = safe_expf(hp[r]);
            Ap[r] = ysq / (Ep[r] + 1e-8f) / nu;
            lw[r] = valid(r) ? th.st_const - 0.5f * hp[r] - np1 * 0.5f * log1pf(Ap[r]) : -3.0e38f;
        }
        float mx = -3.0e38f;
#pragma unroll
        for (int r = 0; r < R; ++r) mx = fmaxf(mx, lw[r]);
        mx = block_max<BS>(mx, red);
        float se[1] = {0.f};
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) se[0] += expf(lw[r] - mx);
        block_sum<BS, 1>(se, red);
        const float inv_se = 1.0f / se[0];
        double tm = 0.0, tr = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            if (i >= n) continue;
            const float wt = expf(lw[r] - mx) * inv_se * glbar;
            const float opa = 1.0f + Ap[r];
            float dh = -0.5f;
            if (hp[r] < 20.0f) dh += 0.5f * np1 * Ap[r] / opa * (Ep[r] / (Ep[r] + 1e-8f));
            hb[r] += wt * dh;
            tn += (double)(wt * (th.dst_const - 0.5f * log1pf(Ap[r]) + 0.5f * np1 * (Ap[r] / nu) / opa));
            // predict
            const float pm = (u[r] > -15.0f && u[r] < 5.0f) ? 1.0f : 0.0f;
            ts -= (double)(mm[r] * isq_m * mb[r] * isig);  // the means' part of the prior's sigma adjoint
            const float mt = mb[r] + hb[r] * pm;
            ts += (double)(hb[r] * pm * noise_z(D, t, i));
            const float ex = safe_expf(xr[r] * 0.5f);
            const float den = ex + 1e-8f;
            const float dlev = -C.gamma * y_prev * (0.5f * ex) / (den * den);
            xb[r] = mt * (rho + dlev);
            vb[r] = vbar[r];
            tm += (double)(mt * (1.0f - rho));
            tr += (double)(mt * (xr[r] - mu));
        }
        tb[0] += tm;
        tb[1] += tr;
        tb[2] += ts;
        tb[3] += tn;
    }
};

}  // namespace svpf
