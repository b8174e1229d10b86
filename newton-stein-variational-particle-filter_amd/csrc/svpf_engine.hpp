// svpf_engine.hpp -- one SVPF filter advanced tick by tick by one workgroup.
//
// The particles of the filter stay in registers (each lane owns R rows) and in
// LDS (a packed {h, s, c, curv} record per particle plus the mixture-component
// means) for every tick the workgroup runs.  A tick follows the reference's
// svpf_step_graph (src/svpf_optimized_graph.cu:915-1389):
//
//   scalars (adaptive mu / sigma, EKF guide, Stein budget)  -> lane 0
//   predict (plain / MIM / guided)                           -> own rows
//   guide (variance preserving or contraction)               -> own rows
//   bandwidth (reference EMA / median pairwise r^2 / Silverman)
//   annealed Stein loop: mixture-prior gradient (O(N^2), one exp per pair)
//                        IMQ transport + Newton + KSD (O(N^2), one rcp per pair)
//   rejuvenation, outputs, host tail (adaptive mu, ksd_prev, ...)
//
// Lane layout: lane t owns rows rg + r*(BS/CS), r < R, with rg = t / CS; the CS
// adjacent lanes of a row split the column loop (j = cs, cs+CS, ...) and meet
// in an xor butterfly.
#pragma once

#include "svpf_device.hpp"

namespace svpf {

enum { MODE_PLAIN = 0, MODE_NEWTON = 1, MODE_FULL = 2 };

struct alignas(16) TickScalars {
    float mu, sigma, y_t, y_prev;
    float ib;            // 1 / bw^2
    float bw, bw_sq;
    float gs;            // guide strength this tick
    float guide_mean, guide_std;
    float ksd;           // KSD of the latest evaluated iteration
    int stop;            // early-stop decision broadcast
    int budget;
    float rs;            // 1 / bw
    float med_est;       // previous tick's median pairwise r^2 (selection seed; 0 = none)
    int pad;
};

// LDS layout for a workgroup running filters of up to nmax particles.
struct LdsLayout {
    int nmax, p2max;
    __host__ __device__ static int pow2_ceil(int n) {
        int p = 1;
        while (p < n) p <<= 1;
        return p;
    }
    // the means / sort buffer: a power of two above n (the median's sorted
    // array is followed by a +inf sentinel)
    __host__ __device__ static int mbuf_len(int n) { return pow2_ceil(n + 1 < 4 ? 4 : n + 1); }
    __host__ __device__ static size_t bytes(int nmax) {
        int p2 = mbuf_len(nmax);
        size_t b = (size_t)nmax * 16;          // hsc records
        b += (size_t)p2 * 4;                   // mixture means / sort buffer
        b += kRedWords * 4;                    // reduction scratch
        b += sizeof(TickScalars);
        return (b + 15) & ~(size_t)15;
    }
};

struct FilterIO {
    const svpf_fparams* P;
    svpf_fstate* S;
    const svpf_fdesc* D;
    float* h;   // pool slices of this filter
    float* hp;
    float* v;
};

// EXT: instantiations with the Heun step and the Student-t state (SVPF_F_HEUN,
// SVPF_F_STUDENT_STATE); without it those flags are not read.
template <int BS, int R, int CS, bool EXT = false>
struct Engine {
    static constexpr int ROWS_PER_PASS = BS / CS;
    static constexpr int CAP = R * ROWS_PER_PASS;

    float4* hsc;       // [n]      {h, s, c, curv}
    float* mbuf;       // [p2]     mixture means (and sort buffer during the bandwidth phase)
    float* red;        // [kRedWords]
    TickScalars* sc;

    int n, p2, rg, cs;
    // per-row registers
    float h[R], hp[R], v[R], lw[R], s[R];

    SVPF_DEV void bind(char* lds, int n_, int nmax) {
        n = __builtin_amdgcn_readfirstlane(n_);  // uniform: loop bounds live in SGPRs
        p2 = LdsLayout::mbuf_len(n);
        int p2max = LdsLayout::mbuf_len(nmax);
        hsc = reinterpret_cast<float4*>(lds);
        mbuf = reinterpret_cast<float*>(lds + (size_t)nmax * 16);
        red = mbuf + p2max;
        sc = reinterpret_cast<TickScalars*>(red + kRedWords);
        rg = threadIdx.x / CS;
        cs = threadIdx.x % CS;
    }

    SVPF_DEV int row(int r) const { return rg + r * ROWS_PER_PASS; }
    SVPF_DEV bool valid(int r) const { return row(r) < n; }

    SVPF_DEV void load(const FilterIO& io) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int i = row(r);
            bool ok = i < n;
            h[r] = ok ? io.h[i] : 0.f;
            hp[r] = ok ? io.hp[i] : 0.f;
            v[r] = ok ? io.v[i] : 0.f;
            lw[r] = 0.f;
            s[r] = 0.f;
        }
    }
    SVPF_DEV void store(const FilterIO& io) {
        if (cs != 0) return;
        float* lwo = io.D->log_w;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int i = row(r);
            if (i < n) {
                io.h[i] = h[r];
                io.hp[i] = hp[r];
                io.v[i] = v[r];
                if (lwo) lwo[i] = lw[r];
            }
        }
    }

    // ---------------------------------------------------------------------
    // Mixture-prior gradient (src/svpf_opt_kernels.cu:354-521), Gaussian state
    // prior: g_prior(h_j) = sum_i r_i (m_i - h_j)/sigma^2 / (sum_i r_i + 1e-8),
    // r_i = exp(-(h_j-m_i)^2/(2 sigma^2) - max)  (:453-470).
    //
    // Evaluated in units scaled by sq = sqrt(log2(e) / (2 sigma^2)): with
    // hh = sq h and mm = sq m, r_i = exp2(dd_min^2 - (hh - mm_i)^2).  The
    // means depend only on the previous tick's particles, so mbuf holds them
    // sorted ascending for the whole Stein loop (sort_means), padded with
    // far-away sentinels to a multiple of 4.  The shift dd_min^2 (the nearest
    // component, the reference's max term) is then a binary search per row
    // (nearest_sq); one pass over the means accumulates: a pair costs
    // sub + fma + exp2 + two accumulates.
    // ---------------------------------------------------------------------

    // Scaled means of the own rows -> mbuf, sorted ascending, sentinels in [n, p2).
    // The transport records are not live between the bandwidth and the first
    // gradient of a tick: their LDS is the sort input.  All threads call it.
    template <class MeanOf>
    SVPF_DEV void sort_means(MeanOf mean_of) {
        float* tmp = reinterpret_cast<float*>(hsc);
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int i = row(r);
                if (i < n) tmp[i] = mean_of(r);
            }
        }
        __syncthreads();
        block_sort<BS>(tmp, mbuf, n);
        for (int i = n + threadIdx.x; i < p2; i += BS) mbuf[i] = 1.0e18f;  // exp2 underflows to 0
        __syncthreads();
    }

    // min_i fl(hh - mm_i)^2 over the sorted means: rounding is monotone, so it
    // is attained at the largest mean below hh or the smallest one at or above
    // it -- bit-identical to the minimum over all means.
    SVPF_DEV float nearest_sq(float hh) const {
        int pos = -1;  // last index with mbuf[pos] < hh
        for (int step = 1 << (31 - __clz(n)); step > 0; step >>= 1) {
            const int q = pos + step;
            pos = (q < n && mbuf[q] < hh) ? q : pos;
        }
        float best = 3.0e38f;
        if (pos >= 0) {
            const float d = hh - mbuf[pos];
            best = d * d;
        }
        if (pos + 1 < n) {
            const float d = hh - mbuf[pos + 1];
            best = fminf(best, d * d);
        }
        return best;
    }

    // SORTED: mbuf holds the means sorted (sort_means) and the shift is a binary
    // search; otherwise (unsorted means) a first pass over the means finds it:
    // sub + mul + min per pair.  OWN (unsorted means): own_mm[r] is the scaled mean
    // of the component row r was predicted from.  Any shift at or above the nearest
    // one gives the same sums up to rounding while no term overflows, and the own
    // component's squared distance is one: its term is 2^0 and every term at most
    // 2^own.  The first pass then runs only in waves with a row whose own component
    // lies more than 10 scaled units away (own > 100).
    template <bool SORTED = true, bool OWN = false>
    SVPF_DEV void prior_sums(float sq, float (&sr)[R], float (&srd)[R], const float* own_mm = nullptr) const {
        float hh[R], off[R];
        bool far = false;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hh[r] = h[r] * sq;
            if (SORTED) {
                off[r] = nearest_sq(hh[r]);
            } else if (OWN) {
                const float d = hh[r] - own_mm[r];
                off[r] = d * d;
                far = far || off[r] > 100.0f;
            } else {
                off[r] = 3.0e38f;
            }
            sr[r] = 0.f;
            srd[r] = 0.f;
        }
        const int nq = (n + 3) >> 2;
        const float4* m4 = reinterpret_cast<const float4*>(mbuf);
        // U loads in flight per step: at one wave per SIMD nothing else hides the LDS latency
        constexpr int U = R >= 2 ? 2 : 4;
        if (!SORTED && (!OWN || __any(far))) {
            if (OWN) {
#pragma unroll
                for (int r = 0; r < R; ++r) off[r] = 3.0e38f;
            }
            auto pass1 = [&](const float4& m) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    float d0 = hh[r] - m.x, d1 = hh[r] - m.y, d2 = hh[r] - m.z, d3 = hh[r] - m.w;
                    off[r] = fminf(off[r], fminf(fminf(d0 * d0, d1 * d1), fminf(d2 * d2, d3 * d3)));
                }
            };
            int q = cs;
            for (; q + (U - 1) * CS < nq; q += U * CS) {
                float4 m[U];
#pragma unroll
                for (int u = 0; u < U; ++u) m[u] = m4[q + u * CS];
#pragma unroll
                for (int u = 0; u < U; ++u) pass1(m[u]);
            }
            for (; q < nq; q += CS) pass1(m4[q]);
#pragma unroll
            for (int r = 0; r < R; ++r) off[r] = group_min<CS>(off[r]);
        }
        auto pass2 = [&](const float4& m) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float d0 = hh[r] - m.x, d1 = hh[r] - m.y, d2 = hh[r] - m.z, d3 = hh[r] - m.w;
                float e0 = __builtin_amdgcn_exp2f(fmaf(-d0, d0, off[r]));
                float e1 = __builtin_amdgcn_exp2f(fmaf(-d1, d1, off[r]));
                float e2 = __builtin_amdgcn_exp2f(fmaf(-d2, d2, off[r]));
                float e3 = __builtin_amdgcn_exp2f(fmaf(-d3, d3, off[r]));
                sr[r] += (e0 + e1) + (e2 + e3);
                srd[r] = fmaf(e0, d0, fmaf(e1, d1, fmaf(e2, d2, fmaf(e3, d3, srd[r]))));
            }
        };
        int q = cs;
        for (; q + (U - 1) * CS < nq; q += U * CS) {
            float4 m[U];
#pragma unroll
            for (int u = 0; u < U; ++u) m[u] = m4[q + u * CS];
#pragma unroll
            for (int u = 0; u < U; ++u) pass2(m[u]);
        }
        for (; q < nq; q += CS) pass2(m4[q]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sr[r] = group_sum<CS>(sr[r]);
            srd[r] = group_sum<CS>(srd[r]);
        }
    }

    // Student-t state prior (src/svpf_opt_kernels.cu:406-440): components
    // log r_i = -a log(1 + d^2/(nu_s sigma^2)), a = (nu_s + 1)/2, shifted by the
    // nearest one.  In units scaled by st = 1/sqrt(nu_s sigma^2) (mbuf holds the
    // means scaled by st, sorted) with q_i = 1 + (hh - mm_i)^2:
    //   r_i = exp2(a (log2 q_min - log2 q_i)),
    //   sr = sum r_i, sg = sum r_i dd_i / q_i, sh = sum r_i (1 - dd_i^2) / q_i^2,
    // so g_prior = -(nu_s+1) st sg/sr and the prior Hessian -(nu_s+1) st^2 sh/sr.
    // A pair costs log2 + exp2 + rcp and five FMA-class operations.
    SVPF_DEV void prior_sums_t(float st, float a, float (&sr)[R], float (&sg)[R], float (&sh)[R]) const {
        float hh[R], base[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hh[r] = h[r] * st;
            base[r] = a * __builtin_amdgcn_logf(1.0f + nearest_sq(hh[r]));
            sr[r] = sg[r] = sh[r] = 0.f;
        }
        const int nq = (n + 3) >> 2;
        const float4* m4 = reinterpret_cast<const float4*>(mbuf);
        auto one = [&](int r, float m) {
            const float d = hh[r] - m;
            const float q = fmaf(d, d, 1.0f);
            const float iq = __builtin_amdgcn_rcpf(q);
            const float e = __builtin_amdgcn_exp2f(fmaf(-a, __builtin_amdgcn_logf(q), base[r]));
            const float u = e * iq;
            sr[r] += e;
            sg[r] = fmaf(u, d, sg[r]);
            sh[r] = fmaf(u * iq, 2.0f - q, sh[r]);
        };
        for (int q = cs; q < nq; q += CS) {
            const float4 m = m4[q];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                one(r, m.x);
                one(r, m.y);
                one(r, m.z);
                one(r, m.w);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sr[r] = group_sum<CS>(sr[r]);
            sg[r] = group_sum<CS>(sg[r]);
            sh[r] = group_sum<CS>(sh[r]);
        }
    }

    // ---------------------------------------------------------------------
    // IMQ Stein transport (src/svpf_opt_kernels.cu:527-599 plain,
    // :742-906 local Newton, :912-1103 full Newton) with the reference KSD
    // (:618-709), in positions scaled by rs = 1/bw: hsc[j] = {h_j rs, s_j,
    // c_j rs, curv_j} with c_j = 2 sgn (local Newton: 2 sgn / curv_j).  Per
    // pair, d = (h_i - h_j) rs, K = 1/(1 + d^2):
    //   acc += K (s_j + c_j rs d K)          = K s_j + c_j (h_i - h_j)/ell^2 K^2
    //   KSD: a1 += K s_j, a2 += d K^2, a3 += s_j d K^2, a4 += K^2 (4 d^2 K - 1)
    //   full Newton: hw += curv_j K + 2/ell^2 K^2 |3 d^2 - 1|, kn += K
    // With d^2 K = 1 - K the two kernel-curvature factors are polynomials in K:
    // 4 d^2 K - 1 = 3 - 4K and K^2 |3 d^2 - 1| = K |3 - 4K|.  On a KSD iteration
    // the transport reuses the KSD sums, acc = a1 + sum_j c_j rs u_j with
    // u = d K^2: 11 VALU operations per pair (one rcp) instead of 15.
    // ---------------------------------------------------------------------
    template <int MODE, bool KSD>
    SVPF_DEV void transport_sums(float ib, float rs, float (&acc)[R], float (&a1)[R], float (&a2)[R],
                                 float (&a3)[R], float (&a4)[R], float (&hw)[R], float (&kn)[R]) const {
        float ht[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ht[r] = h[r] * rs;
            acc[r] = a1[r] = a2[r] = a3[r] = a4[r] = hw[r] = kn[r] = 0.f;
        }
        const float two_ib = 2.0f * ib;
        auto body = [&](const float4& e) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float d = ht[r] - e.x;
                float K = __builtin_amdgcn_rcpf(fmaf(d, d, 1.0f));
                if (KSD) {
                    float K2 = K * K;
                    float u = d * K2;
                    a1[r] = fmaf(K, e.y, a1[r]);
                    a2[r] += u;
                    a3[r] = fmaf(e.y, u, a3[r]);
                    acc[r] = fmaf(e.z, u, acc[r]);  // repulsion; the attraction a1 joins after the loop
                    float t = fmaf(-4.0f, K, 3.0f);
                    a4[r] = fmaf(K2, t, a4[r]);
                    if (MODE == MODE_FULL) {
                        hw[r] = fmaf(e.w, K, hw[r]);
                        hw[r] = fmaf(two_ib, K * fabsf(t), hw[r]);
                        kn[r] += K;
                    }
                } else {
                    float q = fmaf(e.z * d, K, e.y);
                    acc[r] = fmaf(K, q, acc[r]);
                    if (MODE == MODE_FULL) {
                        hw[r] = fmaf(e.w, K, hw[r]);
                        hw[r] = fmaf(two_ib, K * fabsf(fmaf(-4.0f, K, 3.0f)), hw[r]);
                        kn[r] += K;
                    }
                }
            }
        };
        // U loads in flight per step: at one wave per SIMD nothing else hides the LDS latency
        constexpr int U = R >= 4 ? 1 : 8 / R;
        int j = cs;
        for (; j + (U - 1) * CS < n; j += U * CS) {
            float4 e[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = hsc[j + u * CS];
#pragma unroll
            for (int u = 0; u < U; ++u) body(e[u]);
        }
        for (; j < n; j += CS) body(hsc[j]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            acc[r] = group_sum<CS>(acc[r]);
            if (MODE == MODE_FULL) {
                hw[r] = group_sum<CS>(hw[r]);
                kn[r] = group_sum<CS>(kn[r]);
            }
            if (KSD) {
                a1[r] = group_sum<CS>(a1[r]);
                a2[r] = group_sum<CS>(a2[r]);
                a3[r] = group_sum<CS>(a3[r]);
                a4[r] = group_sum<CS>(a4[r]);
                acc[r] += a1[r];
            }
        }
    }

    // ---------------------------------------------------------------------
    // One tick.
    // ---------------------------------------------------------------------
    SVPF_DEV void tick(const FilterIO& io, int t) {
        const svpf_fparams& P = *io.P;
        svpf_fstate& S = *io.S;
        const svpf_fdesc& D = *io.D;
        const uint32_t F = P.flags;
        const bool rng_mode = (D.z_pred == nullptr);
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        const uint32_t tick_id = (uint32_t)S.timestep;
        const float inv_n = 1.0f / (float)n;

        // ---- scalar prologue on lane 0 (src/svpf_optimized_graph.cu:932-1138) ----
        if (threadIdx.x == 0) {
            float y_t = D.y[t];
            float y_prev = S.y_prev;
            float mu = (F & SVPF_F_ADAPT_MU) ? S.mu_state : P.mu;
            float sigma = P.sigma_z;
            if ((F & SVPF_F_ADAPT_SIGMA) && S.timestep > 0) {
                float ve = fmaxf(S.vol_prev, 1e-4f);
                float rz = fabsf(y_t) / ve;
                float boost = 1.0f;
                if (rz > P.sb_thr) boost = 1.0f + (P.sb_max - 1.0f) * fminf((rz - P.sb_thr) / 3.0f, 1.0f);
                sigma = P.sigma_z * boost;
                S.sigma_z_eff = sigma;
            }
            float gs = P.gs_base;
            if (F & SVPF_F_GUIDE) {
                if ((F & SVPF_F_ADAPT_GUIDE) && S.timestep > 0) {
                    float ve = fmaxf(S.vol_prev, 1e-4f);
                    float rz = fabsf(y_t) / ve;
                    float ih = logf(y_t * y_t + 1e-8f) + P.implied_offset;
                    float he = logf(ve * ve + 1e-8f);
                    if (ih - he > 0.0f && rz > P.gs_thr)
                        gs = P.gs_base + (P.gs_max - P.gs_base) * fminf((rz - P.gs_thr) / 3.0f, 1.0f);
                }
                // EKF guide (include/svpf_kernels.cuh:509-541)
                if (!S.guide_initialized) {
                    S.guide_mean = mu;
                    S.guide_var = P.sigma_z * P.sigma_z / (1.0f - P.rho * P.rho);
                    S.guide_initialized = 1;
                }
                float m_pred = mu + P.rho * (S.guide_mean - mu);
                float P_pred = P.rho * P.rho * S.guide_var + P.sigma_z * P.sigma_z;
                float log_y2 = logf(y_t * y_t + 1e-8f);
                float Sv = P_pred + (4.93f + 2.0f);
                float Kg = P_pred / (Sv + 1e-8f);
                float innov = log_y2 - (m_pred - P.implied_offset);
                S.guide_mean = m_pred + Kg * innov;
                S.guide_var = (1.0f - Kg) * P_pred;
            }
            int budget;
            if (S.timestep < 10) budget = P.stein_max;
            else {
                float x = fminf(fmaxf((S.ksd_prev - 0.05f) / (0.50f - 0.05f), 0.0f), 1.0f);
                budget = P.stein_min + (int)(x * (float)(P.stein_max - P.stein_min));
            }
            sc->mu = mu;
            sc->sigma = sigma;
            sc->y_t = y_t;
            sc->y_prev = y_prev;
            sc->gs = gs;
            sc->guide_mean = S.guide_mean;
            sc->guide_std = sqrtf(fmaxf(S.guide_var, 1e-6f));
            sc->budget = budget;
            sc->med_est = S.med_r2;
            sc->stop = 0;
            sc->ksd = 0.f;
        }
        __syncthreads();
        const float mu = sc->mu, sigma = sc->sigma, y_t = sc->y_t, y_prev = sc->y_prev;
        const float h_mean_prev = S.h_mean_prev;

        // ---- predict (src/svpf_opt_kernels.cu:125-276) ----
        {
            const bool mim = (F & (SVPF_F_MIM | SVPF_F_GUIDED)) != 0;
            const bool guided = (F & SVPF_F_GUIDED) != 0;
            const float rho_up = (F & SVPF_F_ASYM_RHO) ? P.rho_up : P.rho;
            const float rho_dn = (F & SVPF_F_ASYM_RHO) ? P.rho_down : P.rho;
            const float d_rho = (F & SVPF_F_LOCAL_PARAMS) ? P.delta_rho : 0.f;
            const float d_sig = (F & SVPF_F_LOCAL_PARAMS) ? P.delta_sigma : 0.f;
            float mean_implied = 0.f;
            if (guided) mean_implied = fmaxf(logf(y_t * y_t + 1e-10f) + P.implied_offset, -5.0f);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int i = row(r);
                if (i >= n) continue;
                float z, u = 1.0f;
                if (rng_mode) {
                    uint32_t o[4];
                    rng.draw((uint32_t)i, tick_id, SVPF_PURPOSE_PREDICT, o);
                    z = normal_from_bits(o[0], o[1]);
                    u = u01_from_bits(o[2]);
                    if (EXT && (F & SVPF_F_STUDENT_STATE) && P.nu_state <= 30.0f) {
                        // Student-t state noise (src/svpf_opt_kernels.cu:31-53): z sqrt(k / chi2_k),
                        // k = max(round(nu_s), 3) squared normals of their own purposes
                        const int k = max((int)(P.nu_state + 0.5f), 3);
                        float chi2 = 0.0f;
                        for (int c = 0; c < k; ++c) {
                            const float w = rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_STATE_T0 + (uint32_t)c);
                            chi2 = fmaf(w, w, chi2);
                        }
                        z *= sqrtf((float)k / (chi2 + 1e-8f));
                    }
                } else {
                    z = D.z_pred[(size_t)t * n + i];
                    if (D.u_pred) u = D.u_pred[(size_t)t * n + i];
                }
                float hi = h[r], hpi = hp[r];
                float lev = P.gamma * y_prev / (safe_expf(hi * 0.5f) + 1e-8f);
                float hn;
                if (mim) {
                    float dev = hi - h_mean_prev;
                    float rho_adj = d_rho * tanhf(dev);
                    float sig_scale = 1.0f + d_sig * fabsf(dev);
                    float scale = (u < P.mim_p) ? P.mim_scale : 1.0f;
                    float br = (hi > hpi) ? rho_up : rho_dn;
                    float rr = fminf(fmaxf(br + rho_adj, 0.0f), 0.999f);
                    float sl = sigma * sig_scale;
                    float mp = mu + rr * (hi - mu) + lev;
                    if (guided) {
                        float zs = (mean_implied - mp) / 2.5f;
                        float act = zs > P.ga_thr ? tanhf(zs - P.ga_thr) : 0.0f;
                        float a = P.ga_base + (P.ga_shock - P.ga_base) * act;
                        mp = (1.0f - a) * mp + a * mean_implied;
                    }
                    hn = clamp_logvol(mp + sl * scale * z);
                } else {
                    hn = clamp_logvol(mu + P.rho * (hi - mu) + sigma * z + lev);
                }
                hp[r] = hi;
                h[r] = hn;
            }
        }
        // predictive log-likelihood terms log p(y_t | h_pred) (loss of the differentiable form)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float vol = safe_expf(h[r]);
            float A = (y_t * y_t) / (vol + 1e-8f) / P.nu;
            lw[r] = valid(r) ? P.st_const - 0.5f * h[r] - (P.nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f)) : -3.0e38f;
        }
        float pred_ll = logmeanexp(lw);

        // ---- guide (src/svpf_opt_kernels.cu:282-348) ----
        if (F & SVPF_F_GUIDE) {
            const float gs = sc->gs, gm = sc->guide_mean;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (F & SVPF_F_GUIDE_PRESERVING) {
                    float nm = (1.0f - gs) * h_mean_prev + gs * gm;
                    h[r] = clamp_logvol(nm + (h[r] - h_mean_prev));
                } else {
                    h[r] = clamp_logvol(h[r] + gs * (gm - h[r]));
                }
            }
        }

        if (D.h_pred && cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) D.h_pred[row(r)] = h[r];
        }

        // ---- bandwidth ----
        bandwidth(P, S, y_t);
        const float ib = sc->ib;
        const float rs = sc->rs;

        // ---- mixture-component means m_i from the previous tick's particles,
        //      scaled by sq ----
        const float inv_s2 = 1.0f / (sigma * sigma);
        const float sq = (EXT && (F & SVPF_F_STUDENT_STATE)) ? sqrtf(inv_s2 / P.nu_state)
                                                    : sqrtf(1.4426950408889634f * 0.5f * inv_s2);
        {
            const float lev_num = P.gamma * y_prev;
            sort_means([&](int r) { return sq * (mu + P.rho * (hp[r] - mu) + lev_num / (expf(hp[r] * 0.5f) + 1e-8f)); });
        }

        // ---- annealed Stein loop (src/svpf_optimized_graph.cu:1057-1317) ----
        const int n_anneal = (F & SVPF_F_ANNEAL) ? max(P.n_anneal, 1) : 1;
        const int budget = sc->budget;
        const float base_step = P.base_step;
        const float temp = (F & SVPF_F_SVLD) ? P.temperature : 0.0f;
        const float sgn2 = (F & SVPF_F_PAPER_SIGN) ? 2.0f : -2.0f;
        const int mode = (F & SVPF_F_NEWTON) ? ((F & SVPF_F_FULL_NEWTON) ? MODE_FULL : MODE_NEWTON) : MODE_PLAIN;
        const bool early = (F & SVPF_F_KSD_EARLY_STOP) != 0;
        int total = 0;
        float ksd = 0.0f, ksd_last = -1.0f;
        bool stop = false;
        for (int ai = 0; ai < n_anneal && !stop; ++ai) {
            float beta;
            if (!(F & SVPF_F_ANNEAL)) beta = 1.0f;
            else if (F & SVPF_F_LINEAR_BETA) beta = n_anneal > 1 ? 0.3f + 0.7f * (float)ai / (float)(n_anneal - 1) : 1.0f;
            else if (F & SVPF_F_ADAPT_BETA) beta = adaptive_beta(S.ksd_prev, ai, n_anneal);
            else beta = (ai % 3 == 0) ? 0.3f : ((ai % 3 == 1) ? 0.65f : 1.0f);
            const float eff = base_step * sqrtf(beta);
            const float dstd = sqrtf(2.0f * eff * temp);
            int si = budget / n_anneal;
            if (ai == n_anneal - 1) si = budget - si * (n_anneal - 1);
            for (int s_it = 0; s_it < si; ++s_it) {
                const bool last = (ai == n_anneal - 1) && (s_it == si - 1);
                const bool want_ksd = last || early;
                const int it_row = total;
                ++total;
                float phi[R], ksdp[R], ho[R];
                if (EXT && (F & SVPF_F_HEUN)) {
                    // Heun predictor-corrector (src/svpf_optimized_graph.cu:1147-1232,
                    // src/svpf_opt_kernels.cu:1293-1566): the operator takes the raw score
                    // (plain, or full Newton's kernel-weighted curvature); predictor from h
                    // with the current RMSProp state and no noise; the corrector below
                    // moves h with the mean of both operators.  KSD: original positions,
                    // scores of the predicted ones.
                    const int hmode = (mode == MODE_FULL) ? MODE_FULL : MODE_PLAIN;
                    float phi1[R], unused[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) ho[r] = h[r];
                    __syncthreads();  // previous transport finished reading hsc
                    gradient(P, beta, y_t, inv_s2, sq, rs, sgn2, hmode);
                    __syncthreads();
                    phi_dispatch(hmode, false, ib, rs, inv_n, phi1, unused);
#pragma unroll
                    for (int r = 0; r < R; ++r) h[r] = clamp_logvol(ho[r] + eff * phi1[r] * rsqrtf(v[r] + P.rms_eps));
                    __syncthreads();
                    gradient(P, beta, y_t, inv_s2, sq, rs, sgn2, hmode);
                    __syncthreads();
                    phi_dispatch(hmode, false, ib, rs, inv_n, phi, unused);
                    if (want_ksd) {
                        __syncthreads();
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            h[r] = ho[r];
                            if (valid(r) && cs == 0) hsc[row(r)] = make_float4(ho[r] * rs, s[r], sgn2 * rs, 1.0f);
                        }
                        __syncthreads();
                        phi_rows<MODE_PLAIN, true>(ib, rs, inv_n, unused, ksdp);
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) phi[r] = 0.5f * (phi1[r] + phi[r]);
                } else {
                    __syncthreads();  // previous transport finished reading hsc; mbuf complete
                    gradient(P, beta, y_t, inv_s2, sq, rs, sgn2, mode);
                    __syncthreads();  // hsc complete
                    phi_dispatch(mode, want_ksd, ib, rs, inv_n, phi, ksdp);
#pragma unroll
                    for (int r = 0; r < R; ++r) ho[r] = h[r];
                }
                // RMSProp + transport + Langevin noise (Jacobi: all phi from the iteration start)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int i = row(r);
                    float p = phi[r];
                    float vn = P.rms_rho * v[r] + (1.0f - P.rms_rho) * p * p;
                    float drift = eff * p * rsqrtf(vn + P.rms_eps);
                    float diff = 0.0f;
                    if (temp > 1e-6f && i < n) {
                        float xi;
                        if (rng_mode) xi = rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_SVLD0 + (uint32_t)it_row);
                        else xi = (it_row < D.xi_rows) ? D.xi[((size_t)t * D.xi_rows + it_row) * n + i] : 0.0f;
                        diff = dstd * xi;
                    }
                    v[r] = vn;
                    h[r] = clamp_logvol(ho[r] + drift + diff);
                }
                if (want_ksd) {
                    float kk[1] = {0.f};
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (valid(r) && cs == 0) kk[0] += ksdp[r];
                    block_sum<BS, 1>(kk, red);
                    float k2 = kk[0] * inv_n * inv_n;
                    ksd = sqrtf(fmaxf(k2, 0.0f));
                    if (early && ksd_last > 0.0f && total >= P.stein_min) {
                        float rel = (ksd_last - ksd) / ksd_last;
                        if (rel < P.ksd_thr) {
                            stop = true;
                            break;
                        }
                    }
                    ksd_last = ksd;
                }
            }
        }

        // ---- rejuvenation (src/svpf_opt_kernels.cu:1248-1275) ----
        if ((F & SVPF_F_REJUV) && (F & SVPF_F_GUIDE) && S.timestep > 10 && S.ksd_prev > P.rej_thr) {
            const float gm = sc->guide_mean, gsd = sc->guide_std;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int i = row(r);
                if (i >= n) continue;
                float u, z;
                if (D.u_rej) u = D.u_rej[(size_t)t * n + i];
                else u = rng.uniform((uint32_t)i, tick_id, SVPF_PURPOSE_REJUV);
                if (u < P.rej_p) {
                    if (D.z_rej) z = D.z_rej[(size_t)t * n + i];
                    else z = rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_REJUV);
                    h[r] = clamp_logvol((1.0f - P.rej_blend) * h[r] + P.rej_blend * (gm + gsd * z));
                }
            }
        }

        // ---- outputs (src/svpf_opt_kernels.cu:1109-1153) ----
        float ll = logmeanexp(lw);
        float sums[2] = {0.f, 0.f};
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) {
                    sums[0] += safe_expf(h[r] * 0.5f);
                    sums[1] += h[r];
                }
        }
        block_sum<BS, 2>(sums, red);
        if (threadIdx.x == 0) {
            float vol = sums[0] * inv_n, hm = sums[1] * inv_n;
            if (D.out) {
                float* o = D.out + (size_t)t * SVPF_OUTW;
                o[SVPF_O_LOGLIK] = ll;
                o[SVPF_O_VOL] = vol;
                o[SVPF_O_HMEAN] = hm;
                o[SVPF_O_KSD] = ksd;
                o[SVPF_O_BW] = sc->bw;
                o[SVPF_O_PRED_LL] = pred_ll;
                o[SVPF_O_STEPS] = (float)total;
                o[SVPF_O_BWSQ] = sc->bw_sq;
            }
            S.last_loglik = ll;
            S.last_vol = vol;
            S.last_h_mean = hm;
            S.last_bw = sc->bw;
            S.last_ksd = ksd;
            S.last_pred_loglik = pred_ll;
            S.last_guide_strength = sc->gs;
            S.stein_steps_used = total;
            // host tail (src/svpf_optimized_graph.cu:1383-1388)
            S.h_mean_prev = hm;
            S.vol_prev = vol;
            S.ksd_prev = ksd;
            if ((F & SVPF_F_ADAPT_MU) && S.timestep > 10) {
                float Pp = S.mu_var + P.mu_q;
                float Rm = P.mu_rscale * sc->bw * sc->bw;
                float Kg = Pp / (Pp + Rm + 1e-8f);
                float mn = S.mu_state + Kg * (hm - S.mu_state);
                S.mu_state = fminf(fmaxf(mn, P.mu_min), P.mu_max);
                S.mu_var = (1.0f - Kg) * Pp;
            }
            S.y_prev = y_t;
            S.timestep += 1;
        }
        __syncthreads();
    }

    // logmeanexp over the valid rows (cs == 0 lanes carry the row values)
    SVPF_DEV float logmeanexp(const float (&x)[R]) {
        float mx = -3.0e38f;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) mx = fmaxf(mx, x[r]);
        mx = block_max<BS>(mx, red);
        float se[1] = {0.f};
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) se[0] += expf(x[r] - mx);
        }
        block_sum<BS, 1>(se, red);
        return mx + logf(fmaxf(se[0] / (float)n, 1e-10f));
    }

    SVPF_DEV static float adaptive_beta(float ksd_prev, int ai, int na) {
        // src/svpf_optimized_graph.cu:1066-1092
        if (ai == na - 1) return 1.0f;
        float beta;
        if (ksd_prev > 0.50f) beta = 0.30f;
        else if (ksd_prev < 0.05f) beta = 0.80f;
        else beta = 0.30f + ((0.50f - ksd_prev) / (0.50f - 0.05f)) * (0.80f - 0.30f);
        beta *= (0.5f + 0.5f * ((float)(ai + 1) / (float)na));
        return fminf(fmaxf(beta, 0.1f), 1.0f);
    }

    // Gradient for own rows; writes the hsc records of own rows.
    // sq: the scale of the means in mbuf (sqrt(log2(e)/(2 sigma^2)) for the
    // Gaussian state prior, 1/sqrt(nu_s sigma^2) for the Student-t one).
    SVPF_DEV void gradient(const svpf_fparams& P, float beta, float y_t, float inv_s2, float sq, float rs, float sgn2,
                           int mode) {
        const uint32_t F = P.flags;
        const bool student = EXT && (F & SVPF_F_STUDENT_STATE) != 0;
        float sr[R], srd[R], shp[R];
        if (student) prior_sums_t(sq, 0.5f * (P.nu_state + 1.0f), sr, srd, shp);
        else prior_sums(sq, sr, srd);
        const float nu = P.nu;
        const float eb = (F & SVPF_F_FAN) ? 1.0f : beta;
        const float ysq = y_t * y_t;
        const float log_y2 = logf(ysq + 1e-10f);
        const float isq = inv_s2 / sq;  // srd is in sq-scaled units
        const float np1s = (P.nu_state + 1.0f) * sq;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float gp, hpr;
            if (student) {
                const float w = 1.0f / (sr[r] + 1e-8f);
                gp = -np1s * srd[r] * w;
                hpr = -np1s * sq * shp[r] * w;
            } else {
                gp = -((srd[r] * isq) / (sr[r] + 1e-8f));
                hpr = -inv_s2;
            }
            float hj = h[r];
            float vol = safe_expf(hj);
            float A = ysq / (vol + 1e-8f) / nu;
            float opa = 1.0f + A;
            lw[r] = (F & SVPF_F_FAN) ? 0.0f : P.st_const - 0.5f * hj - (nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f));
            if (!valid(r)) lw[r] = -3.0e38f;
            float gl;
            if (F & SVPF_F_EXACT_GRAD) gl = -0.5f + 0.5f * (nu + 1.0f) * A / opa - P.lik_offset;
            else gl = (log_y2 - hj + P.lik_offset) / 1.4f;
            float g = fminf(fmaxf(gp + eb * gl, -10.0f), 10.0f);
            float sj = g, cj = sgn2, cv = 1.0f;
            if (mode != MODE_PLAIN) {
                float hl = -0.5f * (nu + 1.0f) * A / (opa * opa);
                cv = fminf(fmaxf(-(hl + hpr), 0.1f), 100.0f);
                if (mode == MODE_NEWTON) {
                    sj = 0.7f * g / cv;
                    cj = sgn2 / cv;
                }
            }
            s[r] = sj;
            int i = row(r);
            if (i < n && cs == 0) hsc[i] = make_float4(hj * rs, sj, cj * rs, cv);
        }
    }

    template <int MODE, bool KSD>
    SVPF_DEV void phi_rows(float ib, float rs, float inv_n, float (&phi)[R], float (&ksdp)[R]) const {
        float acc[R], a1[R], a2[R], a3[R], a4[R], hw[R], kn[R];
        transport_sums<MODE, KSD>(ib, rs, acc, a1, a2, a3, a4, hw, kn);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float p = acc[r] * inv_n;
            if (MODE == MODE_FULL) {
                float hb = fminf(fmaxf(hw[r] / fmaxf(kn[r], 1e-6f), 0.1f), 100.0f);
                p = p / hb * 0.7f;
            }
            phi[r] = p;
            if (KSD) ksdp[r] = s[r] * a1[r] + 2.0f * rs * (s[r] * a2[r] - a3[r]) + 2.0f * ib * a4[r];
            else ksdp[r] = 0.f;
        }
    }

    SVPF_DEV void phi_dispatch(int mode, bool ksd, float ib, float rs, float inv_n, float (&phi)[R],
                               float (&ksdp)[R]) const {
        if (mode == MODE_NEWTON) {
            if (ksd) phi_rows<MODE_NEWTON, true>(ib, rs, inv_n, phi, ksdp);
            else phi_rows<MODE_NEWTON, false>(ib, rs, inv_n, phi, ksdp);
        } else if (mode == MODE_FULL) {
            if (ksd) phi_rows<MODE_FULL, true>(ib, rs, inv_n, phi, ksdp);
            else phi_rows<MODE_FULL, false>(ib, rs, inv_n, phi, ksdp);
        } else {
            if (ksd) phi_rows<MODE_PLAIN, true>(ib, rs, inv_n, phi, ksdp);
            else phi_rows<MODE_PLAIN, false>(ib, rs, inv_n, phi, ksdp);
        }
    }

    // Bandwidth: writes sc->ib, sc->bw, sc->bw_sq.  Uses mbuf as sort buffer.
    SVPF_DEV void bandwidth(const svpf_fparams& P, svpf_fstate& S, float y_t) {
        if (P.bandwidth_mode == SVPF_BW_MEDIAN) {
            // the transport records are not live yet this tick: their LDS is the
            // sort input, then the row brackets and the candidate buffer
            float* tmp = reinterpret_cast<float*>(hsc);
            if (cs == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int i = row(r);
                    if (i < n) tmp[i] = h[r];
                }
            }
            __syncthreads();
            block_sort<BS>(tmp, mbuf, n);
            int* plo = reinterpret_cast<int*>(hsc);
            float med = median_pairwise_r2<BS, 2>(mbuf, n, sc->med_est, plo, plo + n, reinterpret_cast<float*>(plo + 2 * n),
                                                  2 * n, reinterpret_cast<uint32_t*>(red));
            if (threadIdx.x == 0) {
                S.med_r2 = med;
                float bw = fminf(fmaxf(sqrtf(med), 0.01f), 2.0f);
                sc->bw = bw;
                sc->ib = 1.0f / (bw * bw);
                sc->rs = 1.0f / bw;
                sc->bw_sq = med;
            }
            __syncthreads();
            return;
        }
        float sums[2] = {0.f, 0.f};
        float mn = 3.0e38f, mx = -3.0e38f;
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) {
                    sums[0] += h[r];
                    sums[1] += h[r] * h[r];
                    mn = fminf(mn, h[r]);
                    mx = fmaxf(mx, h[r]);
                }
        }
        block_sum<BS, 2>(sums, red);
        mn = block_min<BS>(mn, red);
        mx = block_max<BS>(mx, red);
        if (threadIdx.x == 0) {
            float inv_n = 1.0f / (float)n;
            float mean = sums[0] * inv_n;
            float var = sums[1] * inv_n - mean * mean;
            float bw;
            if (P.bandwidth_mode == SVPF_BW_SILVERMAN) {
                // src/svpf_fully_fused_v2.cu:397-424
                float sd = sqrtf(fmaxf(var, 1e-6f));
                bw = fminf(fmaxf(1.06f * sd * powf((float)n, -0.2f), 0.05f), 2.0f);
                sc->bw_sq = bw * bw;
            } else {
                // src/svpf_opt_kernels.cu:1159-1235
                float spread = mx - mn;
                float bsn = fmaxf(2.0f * var / logf((float)n + 1.0f), 1e-6f);
                float prev = S.bw_sq_ema;
                float bsq = prev > 0.0f ? 0.3f * bsn + 0.7f * prev : bsn;
                float ar = fabsf(y_t);
                float re = S.return_ema, rv = S.return_var;
                re = re > 0.0f ? 0.05f * ar + 0.95f * re : ar;
                rv = rv > 0.0f ? 0.05f * ar * ar + 0.95f * rv : ar * ar;
                S.return_ema = re;
                S.return_var = rv;
                float vr = ar / fmaxf(re, 1e-8f);
                float sf = fminf(spread * 0.5f, 2.0f);
                float comb = fmaxf(vr, sf);
                float scale = fmaxf(fminf(1.0f - 0.25f * fminf(comb - 1.0f, 2.0f), 1.0f), 0.5f);
                bsq *= scale;
                S.bw_sq_ema = bsq;
                bw = fminf(fmaxf(sqrtf(bsq), 0.01f), 2.0f);
                sc->bw_sq = bsq;
            }
            sc->bw = bw;
            sc->ib = 1.0f / (bw * bw);
                sc->rs = 1.0f / bw;
        }
        __syncthreads();
    }
};

}  // namespace svpf
