// svpf_v2.hpp -- the single-workgroup fused step with the semantics of the
// reference's svpf_fully_fused_step_v2 (src/svpf_fully_fused_v2.cu:90-712).
//
// It runs on the single-workgroup engine's pieces (particles in registers, the
// {h, s, c, curv} records and the sorted mixture means in LDS, the one-exp2
// mixture-prior pass and the one-rcp IMQ transport pass) with the fused step's
// own rules:
//   - EKF guide on the adaptive mu, re-initialised when guide_var < 1e-6, the
//     guide strength raised by |y| / vol_prev alone (:180-263);
//   - local parameters whenever delta_rho or delta_sigma > 0, MIM whenever the
//     jump draw falls under mim_jump_prob, guided predict (:270-370);
//   - Silverman bandwidth on a two-pass variance (:397-424);
//   - beta = 0.3 + 0.5 ai / (A - 1) with the last stage at 1, the remainder
//     iterations in the last stage (:430-455);
//   - the prior gradient clamped to +-10 before the likelihood is added
//     (:466-484);
//   - full Newton selected by use_full_newton alone; use_newton scales the
//     repulsion by the stored per-particle term (1/curv, or curv when both are
//     set) (:534-574);
//   - the last iteration's KSD on raw scores, dividing the Newton-scaled ones
//     by 0.7 times the stored term (:579-596);
//   - rejuvenation without the guide requirement, adaptive mu as a clamped EMA
//     (:640-706).
#pragma once

#include "svpf_engine.hpp"

namespace svpf {

struct svpf_v2_kargs {
    float* h;
    float* h_prev;
    float* grad;
    float* logw;
    float* v;
    float* inv_hess;
    float* d_h_mean;
    float* d_vol;
    float* d_loglik;
    float* d_bw;
    float* d_ksd;
    float* d_guide_mean;
    float* d_guide_var;
    float* d_mu_state;
    const float* z_state;
    const float* u_jump;
    const float* z_diff;
    const float* u_rej;
    const float* z_rej;
    uint64_t seed;
    uint32_t stream_id;
    int diff_rows;
    float y_t, y_prev, h_mean_prev, vol_prev, ksd_prev;
    float rho, sigma_z, mu, nu, lik_offset, st_const, implied_offset, gamma;
    float delta_rho, delta_sigma, mim_p, mim_scale;
    float gs_base, gs_max, gs_thr, ga_base, ga_shock, ga_thr;
    float rej_thr, rej_p, rej_blend, mu_alpha, sb_thr, sb_max;
    float step_size, temperature, rms_rho, rms_eps;
    int use_guide, use_guided, use_guide_preserving, use_newton, use_full_newton;
    int use_rejuv, use_adaptive_mu, use_adaptive_sigma;
    int n_stein, n_anneal, sign_mode;
    int n, timestep;
};

template <int BS, int R, int CS>
struct V2Step : Engine<BS, R, CS> {
    using E = Engine<BS, R, CS>;
    using E::cs;
    using E::h;
    using E::hp;
    using E::hsc;
    using E::lw;
    using E::n;
    using E::red;
    using E::row;
    using E::s;
    using E::sc;
    using E::v;
    using E::valid;

    SVPF_DEV void run(const svpf_v2_kargs& A) {
        const bool rng_mode = (A.z_state == nullptr);
        const CounterRng rng{(uint32_t)A.seed, (uint32_t)(A.seed >> 32), A.stream_id};
        const uint32_t tick = (uint32_t)A.timestep;
        const float inv_n = 1.0f / (float)n;

        // ---- phase 0: scalars on lane 0 (:180-263) ----
        if (threadIdx.x == 0) {
            const float mu_eff = A.use_adaptive_mu ? *A.d_mu_state : A.mu;
            float gm_prev = *A.d_guide_mean, gvar = *A.d_guide_var;
            if (gvar < 1e-6f) {
                gm_prev = A.mu;
                gvar = A.sigma_z * A.sigma_z / (1.0f - A.rho * A.rho);
            }
            float boost = 1.0f;
            if (A.use_adaptive_sigma && A.timestep > 0) {
                const float rz = fabsf(A.y_t) / fmaxf(A.vol_prev, 1e-4f);
                if (rz > A.sb_thr) boost = 1.0f + (A.sb_max - 1.0f) * fminf((rz - A.sb_thr) / 3.0f, 1.0f);
            }
            float gmean, gstr;
            if (A.use_guide) {
                const float m_pred = mu_eff + A.rho * (gm_prev - mu_eff);
                const float P_pred = A.rho * A.rho * gvar + A.sigma_z * A.sigma_z;
                const float K = P_pred / (P_pred + (4.93f + 2.0f) + 1e-8f);
                const float innov = logf(A.y_t * A.y_t + 1e-8f) - (m_pred - A.implied_offset);
                gmean = m_pred + K * innov;
                gvar = (1.0f - K) * P_pred;
                *A.d_guide_mean = gmean;
                *A.d_guide_var = gvar;
                const float rz = fabsf(A.y_t) / fmaxf(A.vol_prev, 1e-4f);
                gstr = A.gs_base;
                if (rz > A.gs_thr) gstr = A.gs_base + (A.gs_max - A.gs_base) * fminf((rz - A.gs_thr) / 3.0f, 1.0f);
            } else {
                gmean = A.h_mean_prev;
                gvar = 1.0f;
                gstr = 0.0f;
            }
            sc->mu = mu_eff;
            sc->sigma = A.sigma_z * boost;
            sc->gs = gstr;
            sc->guide_mean = gmean;
            sc->guide_std = sqrtf(fmaxf(gvar, 1e-6f));
        }
        __syncthreads();
        const float mu_eff = sc->mu, sig = sc->sigma, gstr = sc->gs, gmean = sc->guide_mean;
        const float inv_s2 = 1.0f / (sig * sig);
        const float sq = sqrtf(1.4426950408889634f * 0.5f * inv_s2);

        // ---- phase 1: mixture means from the previous posterior, predict (:270-370) ----
        float ho[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            ho[r] = valid(r) ? A.h[i] : 0.0f;
            v[r] = valid(r) ? A.v[i] : 0.0f;
        }
        {
            const float lev_num = A.gamma * A.y_prev;
            E::sort_means([&](int r) { return sq * (mu_eff + A.rho * (ho[r] - mu_eff) + lev_num / (expf(ho[r] * 0.5f) + 1e-8f)); });
        }
        const bool local = A.delta_rho > 0.0f || A.delta_sigma > 0.0f;
        float mean_implied = 0.0f;
        if (A.use_guided) mean_implied = fmaxf(logf(A.y_t * A.y_t + 1e-10f) + A.implied_offset, -5.0f);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            hp[r] = ho[r];
            if (!valid(r)) {
                h[r] = 0.0f;
                continue;
            }
            float lr = A.rho, ls = sig;
            if (local) {
                const float dev = ho[r] - A.h_mean_prev;
                lr = fminf(fmaxf(A.rho + A.delta_rho * tanhf(dev), 0.0f), 0.999f);
                ls = sig * (1.0f + A.delta_sigma * fabsf(dev));
            }
            const float lev = fabsf(A.gamma) > 1e-6f ? A.gamma * A.y_prev / (expf(ho[r] * 0.5f) + 1e-8f) : 0.0f;
            float pm = mu_eff + lr * (ho[r] - mu_eff) + lev;
            float z, u;
            if (rng_mode) {
                uint32_t o[4];
                rng.draw((uint32_t)i, tick, SVPF_PURPOSE_PREDICT, o);
                z = normal_from_bits(o[0], o[1]);
                u = u01_from_bits(o[2]);
            } else {
                z = A.z_state[i];
                u = A.u_jump ? A.u_jump[i] : 1.0f;
            }
            const float ms = ls * (1.0f + (u < A.mim_p ? 1.0f : 0.0f) * (A.mim_scale - 1.0f));
            if (A.use_guided) {
                const float zs = (mean_implied - pm) / 2.5f;
                const float act = zs > A.ga_thr ? tanhf(zs - A.ga_thr) : 0.0f;
                const float al = A.ga_base + (A.ga_shock - A.ga_base) * act;
                pm = (1.0f - al) * pm + al * mean_implied;
            }
            h[r] = clamp_logvol(pm + ms * z);
        }

        // ---- phase 2: guide (:372-390) ----
        if (A.use_guide && gstr > 0.0f) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float x = A.use_guide_preserving ? h[r] + gstr * (gmean - A.h_mean_prev)
                                                       : (1.0f - gstr) * h[r] + gstr * gmean;
                h[r] = clamp_logvol(x);
            }
        }

        // ---- phase 3: Silverman bandwidth, two-pass variance (:397-424) ----
        float sm[1] = {0.0f};
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) sm[0] += h[r];
        }
        block_sum<BS, 1>(sm, red);
        const float mean = sm[0] * inv_n;
        float sv[1] = {0.0f};
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) sv[0] += (h[r] - mean) * (h[r] - mean);
        }
        block_sum<BS, 1>(sv, red);
        const float var = sv[0] * inv_n;
        const float bw = fminf(fmaxf(1.06f * sqrtf(fmaxf(var, 1e-6f)) * powf((float)n, -0.2f), 0.05f), 2.0f);
        const float rs = 1.0f / bw;
        const float ib = rs * rs;

        // ---- phase 4: Stein loop (:430-630) ----
        const int na = A.n_anneal < 1 ? 1 : A.n_anneal;
        const int ns = A.n_stein < 1 ? 1 : A.n_stein;
        const int spa = ns / na, rem = ns - spa * na;
        const float sgn2 = A.sign_mode == 1 ? 2.0f : -2.0f;
        const bool full = A.use_full_newton != 0, newton = A.use_newton != 0;
        const float isq = inv_s2 / sq;
        const float ysq = A.y_t * A.y_t;
        const float np1 = A.nu + 1.0f;
        float g[R], cv[R], ksdp[R];
        int step = 0;
        for (int ai = 0; ai < na; ++ai) {
            const float beta = (na <= 1 || ai == na - 1) ? 1.0f : 0.3f + 0.5f * (float)ai / (float)(na - 1);
            const float eff = A.step_size * sqrtf(beta);
            const int si = spa + (ai == na - 1 ? rem : 0);
            for (int it = 0; it < si; ++it) {
                ++step;
                const bool last = step == ns;
                __syncthreads();  // the previous transport finished reading hsc; means complete
                float sr[R], srd[R];
                E::prior_sums(sq, sr, srd);
                float rawk[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float gp = fminf(fmaxf(-((srd[r] * isq) / (sr[r] + 1e-8f)), -10.0f), 10.0f);
                    const float vol = expf(h[r] * 0.5f);
                    const float Av = ysq / (vol * vol + 1e-8f) / A.nu;
                    const float opa = 1.0f + Av;
                    lw[r] = valid(r) ? A.st_const - 0.5f * h[r] - 0.5f * np1 * logf(opa) : -3.0e38f;
                    const float gl = -0.5f + 0.5f * np1 * Av / opa - A.lik_offset;
                    g[r] = fminf(fmaxf(gp + beta * gl, -10.0f), 10.0f);
                    cv[r] = fminf(fmaxf(-(-0.5f * np1 * Av / (opa * opa) - inv_s2), 0.1f), 100.0f);
                    float sj, hs;
                    if (full) {
                        sj = g[r];
                        hs = cv[r];
                    } else if (newton) {
                        sj = 0.7f * g[r] / cv[r];
                        hs = 1.0f / cv[r];
                    } else {
                        sj = g[r];
                        hs = 1.0f;
                    }
                    rawk[r] = newton ? sj / (0.7f * hs + 1e-8f) : sj;
                    s[r] = g[r];  // own-row score of the KSD: the raw gradient
                    if (valid(r) && cs == 0)
                        hsc[row(r)] = make_float4(h[r] * rs, sj, (newton ? sgn2 * hs : sgn2) * rs, cv[r]);
                }
                __syncthreads();  // records complete
                float phi[R];
                const bool ksd_fused = last && !newton;  // the transport scores are the raw ones
                if (full) {
                    if (ksd_fused) E::template phi_rows<MODE_FULL, true>(ib, rs, inv_n, phi, ksdp);
                    else E::template phi_rows<MODE_FULL, false>(ib, rs, inv_n, phi, ksdp);
                } else {
                    if (ksd_fused) E::template phi_rows<MODE_PLAIN, true>(ib, rs, inv_n, phi, ksdp);
                    else E::template phi_rows<MODE_PLAIN, false>(ib, rs, inv_n, phi, ksdp);
                }
                if (last && newton) {
                    // raw scores for the KSD: records {h rs, raw s_j}
                    __syncthreads();
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (valid(r) && cs == 0) hsc[row(r)] = make_float4(h[r] * rs, rawk[r], 0.0f, 1.0f);
                    __syncthreads();
                    float unused[R];
                    E::template phi_rows<MODE_PLAIN, true>(ib, rs, inv_n, unused, ksdp);
                }
                const float dstd = sqrtf(2.0f * eff * A.temperature);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = row(r);
                    const float p = phi[r];
                    const float vn = A.rms_rho * v[r] + (1.0f - A.rms_rho) * p * p;
                    v[r] = vn;
                    float diff = 0.0f;
                    if (A.temperature > 1e-6f && valid(r)) {
                        float xi;
                        if (rng_mode) xi = rng.normal((uint32_t)i, tick, SVPF_PURPOSE_SVLD0 + (uint32_t)(step - 1));
                        else xi = (A.z_diff && step - 1 < A.diff_rows) ? A.z_diff[(size_t)(step - 1) * n + i] : 0.0f;
                        diff = dstd * xi;
                    }
                    h[r] = clamp_logvol(h[r] + eff * p * rsqrtf(vn + A.rms_eps) + diff);
                }
            }
        }

        // ---- phase 5: rejuvenation (:640-652) ----
        if (A.use_rejuv && A.ksd_prev > A.rej_thr && A.timestep > 10) {
            const float gsd = sc->guide_std;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                if (!valid(r)) continue;
                float u, z = 0.0f;
                if (rng_mode) u = rng.uniform((uint32_t)i, tick, SVPF_PURPOSE_REJUV);
                else u = A.u_rej ? A.u_rej[i] : 1.0f;
                if (u < A.rej_p) {
                    if (rng_mode) z = rng.normal((uint32_t)i, tick, SVPF_PURPOSE_REJUV);
                    else z = A.z_rej ? A.z_rej[i] : 0.0f;
                    h[r] = clamp_logvol((1.0f - A.rej_blend) * h[r] + A.rej_blend * (gmean + gsd * z));
                }
            }
        }

        // ---- phase 6: outputs (:656-707) ----
        float mx = -3.0e38f;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) mx = fmaxf(mx, lw[r]);
        mx = block_max<BS>(mx, red);
        float sums[4] = {0.f, 0.f, 0.f, 0.f};
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) {
                    sums[0] += ksdp[r];
                    sums[1] += h[r];
                    sums[2] += expf(h[r] * 0.5f);
                    sums[3] += expf(lw[r] - mx);
                }
        }
        block_sum<BS, 4>(sums, red);
        if (threadIdx.x == 0) {
            const float k2 = sums[0] * inv_n * inv_n;
            const float hm = sums[1] * inv_n;
            *A.d_ksd = sqrtf(fmaxf(k2, 0.0f));
            *A.d_h_mean = hm;
            *A.d_vol = sums[2] * inv_n;
            *A.d_loglik = mx + logf(fmaxf(sums[3], 1e-10f) * inv_n);
            *A.d_bw = bw;
            if (A.use_adaptive_mu && A.timestep > 10)
                *A.d_mu_state = fminf(fmaxf(A.mu_alpha * hm + (1.0f - A.mu_alpha) * sc->mu, -10.0f), 0.0f);
        }
        if (cs == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                if (!valid(r)) continue;
                A.h[i] = h[r];
                A.h_prev[i] = hp[r];
                A.v[i] = v[r];
                A.grad[i] = g[r];
                A.logw[i] = lw[r];
                if (newton || full) A.inv_hess[i] = cv[r];
            }
        }
    }
};

}  // namespace svpf
