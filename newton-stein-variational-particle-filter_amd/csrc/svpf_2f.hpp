// svpf_2f.hpp -- the coordinate-wise two-factor step, one workgroup per filter.
//
// Log-volatility h = mu + f + s with a fast and a slow mean-zero AR(1) factor
// (src/svpf_optimized_graph.cu:744-897, src/svpf_two_factor.cu:29-468,
// include/svpf_two_factor.cuh:245-300).  A tick:
//
//   lane 0    EKF of the two factors with the innovation split
//   rows      predict (AR(1) per factor, MIM jump on the fast one), guide
//             toward the EKF means (variance preserving or contraction)
//   block     Silverman bandwidth per factor with its floor
//   A x S     gradient (AR(1) prior score per factor plus its sigma^2 share of
//             the shared likelihood score; O(N)), then the 1-D IMQ transport of
//             each factor with its own bandwidth, RMSProp state and Langevin
//             noise: beta = (a+1)/A, step 0.5/S x beta
//   block     outputs: log-mean-exp of the last log-weights, mean exp(h/2),
//             mu + mean f + mean s
//
// The particles stay in registers for every tick of a launch; one 16-byte LDS
// record {f/bw_f, g_f, s/bw_s, g_s} per particle feeds both factors' pair
// sums in one pass over j: per pair and factor sub + fma + rcp + mul + 2 fma.
// Noise comes from the counter generator only.
#pragma once

#include "svpf_engine.hpp"

namespace svpf {

template <int BS, int R, int CS>
struct Engine2F : Engine<BS, R, CS, false> {
    using Base = Engine<BS, R, CS, false>;
    using Base::cs;
    using Base::hsc;
    using Base::n;
    using Base::red;
    using Base::row;
    using Base::sc;
    using Base::valid;
    // fast factor in the base arrays h (f), hp (f_prev), v (RMSProp), s (score);
    // the slow factor here
    float x[R], xp[R], w[R], g[R];

    struct IO {
        const svpf_fparams* P;
        svpf_fstate* S;
        const svpf_fdesc* D;
        float* x2f;  // [6][n]: f | s | f_prev | s_prev | v_f | v_s
        float* h;    // [n]: mu + f + s written at the end of a launch (may be null)
    };

    SVPF_DEV void load2(const IO& io) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            const bool ok = i < n;
            this->h[r] = ok ? io.x2f[i] : 0.f;
            x[r] = ok ? io.x2f[n + i] : 0.f;
            this->hp[r] = ok ? io.x2f[2 * n + i] : 0.f;
            xp[r] = ok ? io.x2f[3 * n + i] : 0.f;
            this->v[r] = ok ? io.x2f[4 * n + i] : 0.f;
            w[r] = ok ? io.x2f[5 * n + i] : 0.f;
            this->lw[r] = 0.f;
            this->s[r] = g[r] = 0.f;
        }
    }

    SVPF_DEV void store2(const IO& io) {
        if (cs != 0) return;
        const float mu = io.P->mu;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            if (i >= n) continue;
            io.x2f[i] = this->h[r];
            io.x2f[n + i] = x[r];
            io.x2f[2 * n + i] = this->hp[r];
            io.x2f[3 * n + i] = xp[r];
            io.x2f[4 * n + i] = this->v[r];
            io.x2f[5 * n + i] = w[r];
            if (io.D->log_w) io.D->log_w[i] = this->lw[r];
            if (io.h) io.h[i] = mu + this->h[r] + x[r];
        }
    }

    // Student-t log density of y at log-variance hh as src/svpf_two_factor.cu:188-196
    // writes it (no clamp on exp(hh)); also the likelihood score.
    SVPF_DEV static float log_w(float hh, float ysq, float nu, float stc, float* ratio_out) {
        const float ratio = ysq / (nu * expf(hh) + 1e-8f);
        *ratio_out = ratio;
        return stc - 0.5f * hh - 0.5f * (nu + 1.0f) * logf(1.0f + ratio);
    }

    SVPF_DEV void tick(const IO& io, int t) {
        const svpf_fparams& P = *io.P;
        svpf_fstate& S = *io.S;
        const svpf_fdesc& D = *io.D;
        const uint32_t F = P.flags;
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        const uint32_t tick_id = (uint32_t)S.timestep;
        const float inv_n = 1.0f / (float)n;
        const float mu = P.mu, nu = P.nu;
        const float rf = P.rho_fast, sf = P.sigma_fast, rs = P.rho_slow, ss = P.sigma_slow;

        // ---- EKF of the factors, innovation split (include/svpf_two_factor.cuh:245-300) ----
        if (threadIdx.x == 0) {
            const float y_t = D.y[t];
            sc->y_t = y_t;
            if (F & SVPF_F_GUIDE) {
                if (!S.g2f_initialized) {
                    S.g2f_mean_fast = 0.0f;
                    S.g2f_mean_slow = 0.0f;
                    S.g2f_var_fast = sf * sf / (1.0f - rf * rf);
                    S.g2f_var_slow = ss * ss / (1.0f - rs * rs);
                    S.g2f_initialized = 1;
                }
                const float mf = rf * S.g2f_mean_fast, ms = rs * S.g2f_mean_slow;
                const float Pf = rf * rf * S.g2f_var_fast + sf * sf;
                const float Ps = rs * rs * S.g2f_var_slow + ss * ss;
                const float innov = logf(y_t * y_t + 1e-8f) + P.implied_offset - (mu + mf + ms);
                const float Sv = Pf + Ps + (4.93f + 2.0f);
                const float Kf = Pf / (Sv + 1e-8f), Ks = Ps / (Sv + 1e-8f);
                S.g2f_mean_fast = mf + Kf * innov;
                S.g2f_mean_slow = ms + Ks * innov;
                S.g2f_var_fast = fmaxf((1.0f - Kf) * Pf, 1e-6f);
                S.g2f_var_slow = fmaxf((1.0f - Ks) * Ps, 1e-6f);
            }
            sc->guide_mean = S.g2f_mean_fast;
            sc->guide_std = S.g2f_mean_slow;  // the slow factor's guide mean (slot reused)
            sc->mu = S.mean_fast_prev;
            sc->sigma = S.mean_slow_prev;
        }
        __syncthreads();
        const float y_t = sc->y_t, ysq = y_t * y_t;
        const float gm_f = sc->guide_mean, gm_s = sc->guide_std;
        const float cm_f = sc->mu, cm_s = sc->sigma;

        // ---- predict (src/svpf_two_factor.cu:55-93) ----
        const bool mim = (F & SVPF_F_MIM) != 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            if (i >= n) continue;
            uint32_t o[4];
            rng.draw((uint32_t)i, tick_id, SVPF_PURPOSE_PREDICT, o);
            float zf = normal_from_bits(o[0], o[1]);
            const float u = u01_from_bits(o[2]);
            const float zs = rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_PREDICT_SLOW);
            if (mim && u < P.mim_p) zf *= P.mim_scale;
            this->hp[r] = this->h[r];
            xp[r] = x[r];
            this->h[r] = clamp_logvol(rf * this->hp[r] + sf * zf);
            x[r] = clamp_logvol(rs * xp[r] + ss * zs);
        }
        // predictive log-likelihood of the predicted particles
        float ratio;
#pragma unroll
        for (int r = 0; r < R; ++r)
            this->lw[r] = valid(r) ? log_w(mu + this->h[r] + x[r], ysq, nu, P.st_const, &ratio) : -3.0e38f;
        const float pred_ll = this->logmeanexp(this->lw);

        // ---- guide (src/svpf_two_factor.cu:320-360), around the previous tick's factor means ----
        if (F & SVPF_F_GUIDE) {
            const float gs = P.guide_strength;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (F & SVPF_F_GUIDE_PRESERVING) {
                    this->h[r] = clamp_logvol((1.0f - gs) * cm_f + gs * gm_f + (this->h[r] - cm_f));
                    x[r] = clamp_logvol((1.0f - gs) * cm_s + gs * gm_s + (x[r] - cm_s));
                } else {
                    this->h[r] = clamp_logvol((1.0f - gs) * this->h[r] + gs * gm_f);
                    x[r] = clamp_logvol((1.0f - gs) * x[r] + gs * gm_s);
                }
            }
        }

        // ---- Silverman bandwidths with floors (src/svpf_two_factor.cu:256-313) ----
        float bwf, bws;
        {
            float m[4] = {0.f, 0.f, 0.f, 0.f};
            if (cs == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (valid(r)) {
                        m[0] += this->h[r];
                        m[1] += this->h[r] * this->h[r];
                        m[2] += x[r];
                        m[3] += x[r] * x[r];
                    }
            }
            block_sum<BS, 4>(m, red);  // every thread receives the sums
            const float nf = powf((float)n, -0.2f);
            const float mf = m[0] * inv_n, ms = m[2] * inv_n;
            const float vf = m[1] * inv_n - mf * mf, vs = m[3] * inv_n - ms * ms;
            bwf = fmaxf(1.06f * sqrtf(fmaxf(vf, 1e-8f)) * nf, P.bw_floor_fast);
            bws = fmaxf(1.06f * sqrtf(fmaxf(vs, 1e-8f)) * nf, P.bw_floor_slow);
        }
        const float rsf = 1.0f / bwf, rss = 1.0f / bws;
        const float sgn2 = (F & SVPF_F_PAPER_SIGN) ? 2.0f : -2.0f;
        const float cf = sgn2 * rsf, csl = sgn2 * rss;

        // ---- Stein loop (src/svpf_optimized_graph.cu:822-868) ----
        const int na = (F & SVPF_F_ANNEAL) ? max(P.n_anneal, 1) : 1;
        const int ns = P.n_stein;
        const float base_step = 0.5f / (float)ns;
        const float temp = (F & SVPF_F_SVLD) ? P.temperature : 0.0f;
        const float s2f = sf * sf + 1e-8f, s2s = ss * ss + 1e-8f;
        const float wf = s2f / (s2f + s2s), wsl = 1.0f - wf;
        const float isf = 1.0f / s2f, iss = 1.0f / s2s;
        const bool exact = (F & SVPF_F_EXACT_GRAD) != 0;
        const float log_y2 = logf(ysq + 1e-8f);
        int it = 0;
        for (int a = 0; a < na; ++a) {
            const float beta = (float)(a + 1) / (float)na;
            const float eff = base_step * beta;
            const float dstd = sqrtf(2.0f * eff * temp);
            for (int k = 0; k < ns; ++k, ++it) {
                // gradient (src/svpf_two_factor.cu:171-249)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float hh = mu + this->h[r] + x[r];
                    const float l = log_w(hh, ysq, nu, P.st_const, &ratio);
                    this->lw[r] = valid(r) ? l : -3.0e38f;
                    const float gl = exact ? -0.5f + 0.5f * (nu + 1.0f) * ratio / (1.0f + ratio + 1e-8f)
                                           : 0.5f * (log_y2 - hh);
                    this->s[r] = -(this->h[r] - rf * this->hp[r]) * isf + beta * wf * gl;
                    g[r] = -(x[r] - rs * xp[r]) * iss + beta * wsl * gl;
                }
                __syncthreads();  // the previous iteration's pair pass has finished reading the records
                if (cs == 0) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (valid(r)) hsc[row(r)] = make_float4(this->h[r] * rsf, this->s[r], x[r] * rss, g[r]);
                }
                __syncthreads();
                float af[R], as[R];
                pair_sums(rsf, rss, cf, csl, af, as);
                // RMSProp + transport + Langevin noise per factor (src/svpf_opt_kernels.cu:527-599)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = row(r);
                    const float pf = af[r] * inv_n, ps = as[r] * inv_n;
                    const float vfn = P.rms_rho * this->v[r] + (1.0f - P.rms_rho) * pf * pf;
                    const float vsn = P.rms_rho * w[r] + (1.0f - P.rms_rho) * ps * ps;
                    float df = 0.0f, ds = 0.0f;
                    if (temp > 1e-6f && i < n) {
                        df = dstd * rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_SVLD0 + (uint32_t)it);
                        ds = dstd * rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_SVLD_SLOW0 + (uint32_t)it);
                    }
                    this->v[r] = vfn;
                    w[r] = vsn;
                    this->h[r] = clamp_logvol(this->h[r] + eff * pf * rsqrtf(vfn + P.rms_eps) + df);
                    x[r] = clamp_logvol(x[r] + eff * ps * rsqrtf(vsn + P.rms_eps) + ds);
                }
            }
        }

        // ---- outputs (src/svpf_two_factor.cu:394-468) ----
        const float ll = this->logmeanexp(this->lw);
        float sums[3] = {0.f, 0.f, 0.f};
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) {
                    sums[0] += this->h[r];
                    sums[1] += x[r];
                    sums[2] += expf(0.5f * (mu + this->h[r] + x[r]));
                }
        }
        block_sum<BS, 3>(sums, red);
        if (threadIdx.x == 0) {
            const float mfast = sums[0] * inv_n, mslow = sums[1] * inv_n, vol = sums[2] * inv_n;
            const float hm = mu + mfast + mslow;
            if (D.out) {
                float* o = D.out + (size_t)t * SVPF_OUTW;
                o[SVPF_O_LOGLIK] = ll;
                o[SVPF_O_VOL] = vol;
                o[SVPF_O_HMEAN] = hm;
                o[SVPF_O_KSD] = 0.0f;
                o[SVPF_O_BW] = bwf;
                o[SVPF_O_PRED_LL] = pred_ll;
                o[SVPF_O_STEPS] = (float)it;
                o[SVPF_O_BWSQ] = bws;
            }
            S.last_loglik = ll;
            S.last_vol = vol;
            S.last_h_mean = hm;
            S.last_bw = bwf;
            S.last_ksd = 0.0f;
            S.last_pred_loglik = pred_ll;
            S.last_guide_strength = P.guide_strength;
            S.stein_steps_used = it;
            S.mean_fast_prev = mfast;
            S.mean_slow_prev = mslow;
            S.h_mean_prev = hm;
            S.vol_prev = vol;
            S.y_prev = y_t;
            S.timestep += 1;
        }
        __syncthreads();
    }

    // Both factors' IMQ operator sums for the own rows: af = sum_j K (g_j + c d K)
    // over the fast coordinate, as over the slow one, d in bandwidth units.
    SVPF_DEV void pair_sums(float rsf, float rss, float cf, float csl, float (&af)[R], float (&as)[R]) const {
        float tf[R], ts[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            tf[r] = this->h[r] * rsf;
            ts[r] = x[r] * rss;
            af[r] = as[r] = 0.f;
        }
        auto body = [&](const float4& e) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float d1 = tf[r] - e.x;
                const float k1 = __builtin_amdgcn_rcpf(fmaf(d1, d1, 1.0f));
                af[r] = fmaf(k1, fmaf(cf * d1, k1, e.y), af[r]);
                const float d2 = ts[r] - e.z;
                const float k2 = __builtin_amdgcn_rcpf(fmaf(d2, d2, 1.0f));
                as[r] = fmaf(k2, fmaf(csl * d2, k2, e.w), as[r]);
            }
        };
        constexpr int U = R >= 4 ? 1 : 4 / R;
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
            af[r] = group_sum<CS>(af[r]);
            as[r] = group_sum<CS>(as[r]);
        }
    }
};

}  // namespace svpf
