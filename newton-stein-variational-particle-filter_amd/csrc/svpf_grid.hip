// svpf_grid.hip -- the SVPF tick for one large filter (n up to 16384) as a
// sequence of grid-wide kernels.
//
// The single-workgroup engine keeps a whole filter in one workgroup's LDS,
// which bounds n near 4096.  Here a tick follows the same stages
// (src/svpf_optimized_graph.cu:915-1389) over the whole GPU:
//
//   grid_prologue    1 workgroup   scalar prologue: adaptive mu / sigma, EKF guide, Stein budget
//   grid_predict     n/256         predict (plain / MIM / guided), predictive log weights, guide,
//                                  block partials for the reductions
//   grid_bandwidth   1 x 1024      predictive loglik; bandwidth (median pairwise r^2 by an LDS sort and
//                                  the exact selection with 16-bit row brackets, or the EMA / Silverman rules)
//   grid_means       n/256         scaled mixture-component means
//   per Stein iteration k < max:
//     grid_gradient  G x 256       mixture prior (component means streamed through LDS tiles) + likelihood
//                                  + Newton curvature -> transport records {h/bw, s, c/bw, curv}
//     grid_transport G x 256       IMQ transport over record tiles in LDS, RMSProp, Langevin noise; KSD
//                                  partials when the iteration needs them
//     grid_ksd       1 workgroup   KSD, early-stop decision
//   grid_outputs     n/256         rejuvenation, output partials
//   grid_finish      1 workgroup   outputs and the host tail
//
// The Stein budget and early stop are decided on the device: iterations past
// them return at once, so a tick is a fixed launch sequence (capturable as one
// graph).  Rows of the O(N^2) kernels are spread CS lanes per row so that G
// covers the CUs; the per-particle math is the engine's (svpf_engine.hpp).
#include <hip/hip_runtime.h>

#include "svpf_engine.hpp"
#include "svpf_kernels.h"

using namespace svpf;

namespace {

constexpr int kGB = 256;          // block size of the grid kernels
constexpr int kMeanTile = 4096;   // mixture means per LDS tile (16 KiB)
constexpr int kRecTile = 2048;    // transport records per LDS tile (32 KiB)
constexpr int kMaxN = 16384;      // largest filter of this path
constexpr int kMaxG = 4096;       // most row workgroups of one launch (n / 4 at 64 lanes per row)
enum { GP_LWMAX = 0, GP_LWSUM = 1, GP_H = 2, GP_H2 = 3, GP_HMIN = 4, GP_HMAX = 5, GP_VOL = 6, GP_KSD = 7, GP_N = 8 };

struct GridScalars {
    float mu, sigma, y_t, y_prev;
    float gs, guide_mean, guide_std, ib;
    float bw, bw_sq, rs, sq;
    float inv_s2, pred_ll, ksd, ksd_last;
    int budget, stop, total, pad;
};

// beta and step of global Stein iteration it under the budget (the engine's loop order)
SVPF_DEV void iter_schedule(const svpf_fparams& P, const svpf_fstate& S, int budget, int it, float& beta, bool& last) {
    const uint32_t F = P.flags;
    const int na = (F & SVPF_F_ANNEAL) ? max(P.n_anneal, 1) : 1;
    const int per = budget / na;
    int ai = 0, start = 0;
    for (; ai < na; ++ai) {
        const int si = (ai == na - 1) ? budget - per * (na - 1) : per;
        if (it < start + si) break;
        start += si;
    }
    if (ai >= na) ai = na - 1;
    last = (it == budget - 1);
    if (!(F & SVPF_F_ANNEAL)) beta = 1.0f;
    else if (F & SVPF_F_LINEAR_BETA) beta = na > 1 ? 0.3f + 0.7f * (float)ai / (float)(na - 1) : 1.0f;
    else if (F & SVPF_F_ADAPT_BETA) beta = Engine<256, 1, 1>::adaptive_beta(S.ksd_prev, ai, na);
    else beta = (ai % 3 == 0) ? 0.3f : ((ai % 3 == 1) ? 0.65f : 1.0f);
}

__global__ void grid_prologue(const svpf_fparams* __restrict__ Pp, svpf_fstate* __restrict__ Sp,
                              const svpf_fdesc* __restrict__ Dp, int t, GridScalars* __restrict__ g) {
    if (threadIdx.x != 0) return;
    const svpf_fparams& P = *Pp;
    svpf_fstate& S = *Sp;
    const uint32_t F = P.flags;
    const float y_t = Dp->y[t];
    const float mu = (F & SVPF_F_ADAPT_MU) ? S.mu_state : P.mu;
    float sigma = P.sigma_z;
    if ((F & SVPF_F_ADAPT_SIGMA) && S.timestep > 0) {
        const float ve = fmaxf(S.vol_prev, 1e-4f), rz = fabsf(y_t) / ve;
        float boost = 1.0f;
        if (rz > P.sb_thr) boost = 1.0f + (P.sb_max - 1.0f) * fminf((rz - P.sb_thr) / 3.0f, 1.0f);
        sigma = P.sigma_z * boost;
        S.sigma_z_eff = sigma;
    }
    float gs = P.gs_base;
    if (F & SVPF_F_GUIDE) {
        if ((F & SVPF_F_ADAPT_GUIDE) && S.timestep > 0) {
            const float ve = fmaxf(S.vol_prev, 1e-4f), rz = fabsf(y_t) / ve;
            const float ih = logf(y_t * y_t + 1e-8f) + P.implied_offset, he = logf(ve * ve + 1e-8f);
            if (ih - he > 0.0f && rz > P.gs_thr) gs = P.gs_base + (P.gs_max - P.gs_base) * fminf((rz - P.gs_thr) / 3.0f, 1.0f);
        }
        if (!S.guide_initialized) {
            S.guide_mean = mu;
            S.guide_var = P.sigma_z * P.sigma_z / (1.0f - P.rho * P.rho);
            S.guide_initialized = 1;
        }
        // include/svpf_kernels.cuh:509-541
        const float m_pred = mu + P.rho * (S.guide_mean - mu);
        const float P_pred = P.rho * P.rho * S.guide_var + P.sigma_z * P.sigma_z;
        const float Sv = P_pred + (4.93f + 2.0f);
        const float Kg = P_pred / (Sv + 1e-8f);
        const float innov = logf(y_t * y_t + 1e-8f) - (m_pred - P.implied_offset);
        S.guide_mean = m_pred + Kg * innov;
        S.guide_var = (1.0f - Kg) * P_pred;
    }
    int budget;
    if (S.timestep < 10) budget = P.stein_max;
    else {
        const float x = fminf(fmaxf((S.ksd_prev - 0.05f) / (0.50f - 0.05f), 0.0f), 1.0f);
        budget = P.stein_min + (int)(x * (float)(P.stein_max - P.stein_min));
    }
    g->mu = mu;
    g->sigma = sigma;
    g->y_t = y_t;
    g->y_prev = S.y_prev;
    g->gs = gs;
    g->guide_mean = S.guide_mean;
    g->guide_std = sqrtf(fmaxf(S.guide_var, 1e-6f));
    g->inv_s2 = 1.0f / (sigma * sigma);
    g->sq = sqrtf(1.4426950408889634f * 0.5f * g->inv_s2);
    g->budget = budget;
    g->stop = 0;
    g->total = 0;
    g->ksd = 0.0f;
    g->ksd_last = -1.0f;
}

// workgroup reduction of GP_N partials written by thread 0 to part[blockIdx.x]
SVPF_DEV void block_partials(float (&v)[GP_N], float* red, float* part) {
    float s4[4] = {v[GP_LWSUM], v[GP_H], v[GP_H2], v[GP_VOL]};
    const float mx = block_max<kGB>(v[GP_LWMAX], red);
    // the log-weight sum is relative to each thread's own max: rescale to the block max
    s4[0] = v[GP_LWMAX] > -1.0e38f ? s4[0] * expf(v[GP_LWMAX] - mx) : 0.0f;
    block_sum<kGB, 4>(s4, red);
    float k1[1] = {v[GP_KSD]};
    block_sum<kGB, 1>(k1, red);
    const float mn = block_min<kGB>(v[GP_HMIN], red);
    const float mxh = block_max<kGB>(v[GP_HMAX], red);
    if (threadIdx.x == 0) {
        float* o = part + (size_t)blockIdx.x * GP_N;
        o[GP_LWMAX] = mx;
        o[GP_LWSUM] = s4[0];
        o[GP_H] = s4[1];
        o[GP_H2] = s4[2];
        o[GP_VOL] = s4[3];
        o[GP_KSD] = k1[0];
        o[GP_HMIN] = mn;
        o[GP_HMAX] = mxh;
    }
}

__global__ __launch_bounds__(kGB) void grid_predict(const svpf_fparams* __restrict__ Pp,
                                                    const svpf_fstate* __restrict__ Sp,
                                                    const svpf_fdesc* __restrict__ Dp, int t,
                                                    const GridScalars* __restrict__ g, float* __restrict__ h,
                                                    float* __restrict__ hp, float* __restrict__ lw,
                                                    float* __restrict__ part, int n) {
    __shared__ float red[64];
    const svpf_fparams& P = *Pp;
    const svpf_fstate& S = *Sp;
    const svpf_fdesc& D = *Dp;
    const uint32_t F = P.flags;
    const int i = blockIdx.x * kGB + threadIdx.x;
    const float mu = g->mu, sigma = g->sigma, y_t = g->y_t, y_prev = g->y_prev;
    float v[GP_N] = {-3.0e38f, 0.f, 0.f, 0.f, 3.0e38f, -3.0e38f, 0.f, 0.f};
    if (i < n) {
        const bool mim = (F & (SVPF_F_MIM | SVPF_F_GUIDED)) != 0, guided = (F & SVPF_F_GUIDED) != 0;
        const float rho_up = (F & SVPF_F_ASYM_RHO) ? P.rho_up : P.rho;
        const float rho_dn = (F & SVPF_F_ASYM_RHO) ? P.rho_down : P.rho;
        const float d_rho = (F & SVPF_F_LOCAL_PARAMS) ? P.delta_rho : 0.f;
        const float d_sig = (F & SVPF_F_LOCAL_PARAMS) ? P.delta_sigma : 0.f;
        float z, u = 1.0f;
        if (D.z_pred == nullptr) {
            CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
            uint32_t o[4];
            rng.draw((uint32_t)i, (uint32_t)S.timestep, SVPF_PURPOSE_PREDICT, o);
            z = normal_from_bits(o[0], o[1]);
            u = u01_from_bits(o[2]);
        } else {
            z = D.z_pred[(size_t)t * n + i];
            if (D.u_pred) u = D.u_pred[(size_t)t * n + i];
        }
        const float hi = h[i], hpi = hp[i];
        const float lev = P.gamma * y_prev / (safe_expf(hi * 0.5f) + 1e-8f);
        float hn;
        if (mim) {
            const float dev = hi - S.h_mean_prev;
            const float rho_adj = d_rho * tanhf(dev), sig_scale = 1.0f + d_sig * fabsf(dev);
            const float scale = (u < P.mim_p) ? P.mim_scale : 1.0f;
            const float br = (hi > hpi) ? rho_up : rho_dn;
            const float rr = fminf(fmaxf(br + rho_adj, 0.0f), 0.999f);
            float mp = mu + rr * (hi - mu) + lev;
            if (guided) {
                const float mi = fmaxf(logf(y_t * y_t + 1e-10f) + P.implied_offset, -5.0f);
                const float zs = (mi - mp) / 2.5f;
                const float act = zs > P.ga_thr ? tanhf(zs - P.ga_thr) : 0.0f;
                const float a = P.ga_base + (P.ga_shock - P.ga_base) * act;
                mp = (1.0f - a) * mp + a * mi;
            }
            hn = clamp_logvol(mp + sigma * sig_scale * scale * z);
        } else {
            hn = clamp_logvol(mu + P.rho * (hi - mu) + sigma * z + lev);
        }
        const float A = (y_t * y_t) / (safe_expf(hn) + 1e-8f) / P.nu;
        const float l = P.st_const - 0.5f * hn - (P.nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f));
        if (F & SVPF_F_GUIDE) {
            const float gs = g->gs, gm = g->guide_mean, hmp = S.h_mean_prev;
            if (F & SVPF_F_GUIDE_PRESERVING) hn = clamp_logvol((1.0f - gs) * hmp + gs * gm + (hn - hmp));
            else hn = clamp_logvol(hn + gs * (gm - hn));
        }
        hp[i] = hi;
        h[i] = hn;
        lw[i] = l;
        if (D.h_pred) D.h_pred[i] = hn;
        v[GP_LWMAX] = l;
        v[GP_LWSUM] = 1.0f;
        v[GP_H] = hn;
        v[GP_H2] = hn * hn;
        v[GP_HMIN] = hn;
        v[GP_HMAX] = hn;
    }
    block_partials(v, red, part);
}

// sort of n floats and the exact median selection with 16-bit row brackets
__global__ __launch_bounds__(1024) void grid_bandwidth(const svpf_fparams* __restrict__ Pp,
                                                       svpf_fstate* __restrict__ Sp, GridScalars* __restrict__ g,
                                                       const float* __restrict__ h, const float* __restrict__ part,
                                                       int nb, int n) {
    extern __shared__ __align__(16) char lds[];
    constexpr int BS = 1024;
    __shared__ float red[80];
    const svpf_fparams& P = *Pp;
    svpf_fstate& S = *Sp;
    // predictive loglik and moments from the predict partials
    float M = -3.0e38f;
    for (int b = threadIdx.x; b < nb; b += BS) M = fmaxf(M, part[(size_t)b * GP_N + GP_LWMAX]);
    M = block_max<BS>(M, red);
    float sums[4] = {0.f, 0.f, 0.f, 0.f};
    float mn = 3.0e38f, mx = -3.0e38f;
    for (int b = threadIdx.x; b < nb; b += BS) {
        const float* p = part + (size_t)b * GP_N;
        if (p[GP_LWMAX] > -1.0e38f) sums[0] += p[GP_LWSUM] * expf(p[GP_LWMAX] - M);
        sums[1] += p[GP_H];
        sums[2] += p[GP_H2];
        mn = fminf(mn, p[GP_HMIN]);
        mx = fmaxf(mx, p[GP_HMAX]);
    }
    block_sum<BS, 4>(sums, red);
    mn = block_min<BS>(mn, red);
    mx = block_max<BS>(mx, red);
    const float inv_n = 1.0f / (float)n;
    float bw, bsq;
    if (P.bandwidth_mode == SVPF_BW_MEDIAN) {
        int ps = 128;
        while (ps < n) ps <<= 1;
        float* src = reinterpret_cast<float*>(lds);            // ps floats; then the 16-bit brackets
        float* dst = src + ps;                                  // ps + 4 floats (sorted, +inf sentinel)
        float* cand = dst + ps + 4;                             // kMedianCands floats
        for (int i = threadIdx.x; i < n; i += BS) src[i] = h[i];
        __syncthreads();
        block_sort<BS>(src, dst, n);
        uint16_t* plo = reinterpret_cast<uint16_t*>(src);
        const float med = median_pairwise_r2<BS, 2, uint16_t>(dst, n, S.med_r2, plo, plo + n, cand,
                                                              min(kMedianCands, 2 * n), reinterpret_cast<uint32_t*>(red));
        bw = fminf(fmaxf(sqrtf(med), 0.01f), 2.0f);
        bsq = med;
        if (threadIdx.x == 0) S.med_r2 = med;
    } else {
        const float mean = sums[1] * inv_n, var = sums[2] * inv_n - mean * mean;
        if (P.bandwidth_mode == SVPF_BW_SILVERMAN) {
            bw = fminf(fmaxf(1.06f * sqrtf(fmaxf(var, 1e-6f)) * powf((float)n, -0.2f), 0.05f), 2.0f);
            bsq = bw * bw;
        } else {  // src/svpf_opt_kernels.cu:1159-1235
            const float y_t = g->y_t;
            const float bsn = fmaxf(2.0f * var / logf((float)n + 1.0f), 1e-6f);
            const float prev = S.bw_sq_ema;
            float b2 = prev > 0.0f ? 0.3f * bsn + 0.7f * prev : bsn;
            const float ar = fabsf(y_t);
            float re = S.return_ema, rv = S.return_var;
            re = re > 0.0f ? 0.05f * ar + 0.95f * re : ar;
            rv = rv > 0.0f ? 0.05f * ar * ar + 0.95f * rv : ar * ar;
            const float vr = ar / fmaxf(re, 1e-8f);
            const float sf = fminf((mx - mn) * 0.5f, 2.0f);
            const float comb = fmaxf(vr, sf);
            b2 *= fmaxf(fminf(1.0f - 0.25f * fminf(comb - 1.0f, 2.0f), 1.0f), 0.5f);
            if (threadIdx.x == 0) {
                S.return_ema = re;
                S.return_var = rv;
                S.bw_sq_ema = b2;
            }
            bw = fminf(fmaxf(sqrtf(b2), 0.01f), 2.0f);
            bsq = b2;
        }
    }
    if (threadIdx.x == 0) {
        g->pred_ll = M + logf(fmaxf(sums[0] / (float)n, 1e-10f));
        g->bw = bw;
        g->bw_sq = bsq;
        g->ib = 1.0f / (bw * bw);
        g->rs = 1.0f / bw;
    }
}

__global__ __launch_bounds__(kGB) void grid_means(const svpf_fparams* __restrict__ Pp,
                                                  const GridScalars* __restrict__ g, const float* __restrict__ hp,
                                                  float* __restrict__ mm, int n, int npad) {
    const int i = blockIdx.x * kGB + threadIdx.x;
    if (i >= npad) return;
    if (i >= n) {
        mm[i] = 1.0e18f;  // sentinel: exp2 underflows to 0
        return;
    }
    const svpf_fparams& P = *Pp;
    const float mu = g->mu;
    mm[i] = g->sq * (mu + P.rho * (hp[i] - mu) + P.gamma * g->y_prev / (expf(hp[i] * 0.5f) + 1e-8f));
}

// gradient of rows r = blockIdx.x * (kGB/CS) + threadIdx.x / CS; CS lanes split the components
template <int CS>
__global__ __launch_bounds__(kGB) void grid_gradient(const svpf_fparams* __restrict__ Pp,
                                                     const svpf_fstate* __restrict__ Sp,
                                                     const GridScalars* __restrict__ g, const float* __restrict__ h,
                                                     const float* __restrict__ mm, float4* __restrict__ rec,
                                                     float* __restrict__ lw, int it, int n, int npad) {
    __shared__ __align__(16) float tile[kMeanTile];
    if (it >= g->budget || g->stop) return;  // uniform across the grid
    const svpf_fparams& P = *Pp;
    const uint32_t F = P.flags;
    const int i = blockIdx.x * (kGB / CS) + threadIdx.x / CS, cs = threadIdx.x % CS;
    float beta;
    bool last;
    iter_schedule(P, *Sp, g->budget, it, beta, last);
    const float sq = g->sq, inv_s2 = g->inv_s2, rs = g->rs;
    const float hi = i < n ? h[i] : 0.f;
    const float hh = hi * sq;
    // pass 1: nearest component (the reference's max shift); pass 2: sums
    float off = 3.0e38f, sr = 0.f, srd = 0.f;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) off = group_min<CS>(off);
        for (int j0 = 0; j0 < npad; j0 += kMeanTile) {
            const int len = min(kMeanTile, npad - j0);
            __syncthreads();
            for (int k = threadIdx.x * 4; k < len; k += kGB * 4)
                *reinterpret_cast<float4*>(tile + k) = *reinterpret_cast<const float4*>(mm + j0 + k);
            __syncthreads();
            const float4* t4 = reinterpret_cast<const float4*>(tile);
            for (int q = cs; q < (len >> 2); q += CS) {
                const float4 m = t4[q];
                const float d0 = hh - m.x, d1 = hh - m.y, d2 = hh - m.z, d3 = hh - m.w;
                if (pass == 0) {
                    off = fminf(off, fminf(fminf(d0 * d0, d1 * d1), fminf(d2 * d2, d3 * d3)));
                } else {
                    const float e0 = __builtin_amdgcn_exp2f(fmaf(-d0, d0, off));
                    const float e1 = __builtin_amdgcn_exp2f(fmaf(-d1, d1, off));
                    const float e2 = __builtin_amdgcn_exp2f(fmaf(-d2, d2, off));
                    const float e3 = __builtin_amdgcn_exp2f(fmaf(-d3, d3, off));
                    sr += (e0 + e1) + (e2 + e3);
                    srd = fmaf(e0, d0, fmaf(e1, d1, fmaf(e2, d2, fmaf(e3, d3, srd))));
                }
            }
        }
    }
    sr = group_sum<CS>(sr);
    srd = group_sum<CS>(srd);
    if (i >= n || cs != 0) return;
    const float nu = P.nu, ysq = g->y_t * g->y_t;
    const float gp = -((srd * (inv_s2 / sq)) / (sr + 1e-8f));
    const float A = ysq / (safe_expf(hi) + 1e-8f) / nu, opa = 1.0f + A;
    lw[i] = (F & SVPF_F_FAN) ? 0.0f : P.st_const - 0.5f * hi - (nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f));
    const float gl = (F & SVPF_F_EXACT_GRAD) ? -0.5f + 0.5f * (nu + 1.0f) * A / opa - P.lik_offset
                                              : (logf(ysq + 1e-10f) - hi + P.lik_offset) / 1.4f;
    const float eb = (F & SVPF_F_FAN) ? 1.0f : beta;
    const float gg = fminf(fmaxf(gp + eb * gl, -10.0f), 10.0f);
    const float sgn2 = (F & SVPF_F_PAPER_SIGN) ? 2.0f : -2.0f;
    float sj = gg, cj = sgn2, cv = 1.0f;
    if (F & SVPF_F_NEWTON) {
        cv = fminf(fmaxf(0.5f * (nu + 1.0f) * A / (opa * opa) + inv_s2, 0.1f), 100.0f);
        if (!(F & SVPF_F_FULL_NEWTON)) {
            sj = 0.7f * gg / cv;
            cj = sgn2 / cv;
        }
    }
    rec[i] = make_float4(hi * rs, sj, cj * rs, cv);
}

template <int CS, int MODE, bool KSD>
SVPF_DEV void grid_transport_rows(const svpf_fparams& P, const svpf_fstate& S, const svpf_fdesc& D,
                                  const GridScalars* g, const float4* __restrict__ rec, float* __restrict__ h,
                                  float* __restrict__ v, float* __restrict__ part, float4* tile, int it, int t, int n) {
    const int i = blockIdx.x * (kGB / CS) + threadIdx.x / CS, cs = threadIdx.x % CS;
    const float ib = g->ib, rs = g->rs, two_ib = 2.0f * ib;
    const float4 own = i < n ? rec[i] : make_float4(0.f, 0.f, 0.f, 1.f);
    const float ht = own.x;
    float acc = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f, hw = 0.f, kn = 0.f;
    for (int j0 = 0; j0 < n; j0 += kRecTile) {
        const int len = min(kRecTile, n - j0);
        __syncthreads();
        for (int k = threadIdx.x; k < len; k += kGB) tile[k] = rec[j0 + k];
        __syncthreads();
        for (int j = cs; j < len; j += CS) {
            const float4 e = tile[j];
            const float d = ht - e.x;
            const float K = __builtin_amdgcn_rcpf(fmaf(d, d, 1.0f));
            acc = fmaf(K, fmaf(e.z * d, K, e.y), acc);
            if (MODE == MODE_FULL) {
                const float K2 = K * K;
                hw = fmaf(e.w, K, hw);
                hw = fmaf(two_ib * K2, fabsf(fmaf(3.0f * d, d, -1.0f)), hw);
                kn += K;
            }
            if (KSD) {
                const float K2 = K * K, u = d * K2;
                a1 = fmaf(K, e.y, a1);
                a2 += u;
                a3 = fmaf(e.y, u, a3);
                a4 = fmaf(K2, fmaf(4.0f * d * d, K, -1.0f), a4);
            }
        }
    }
    acc = group_sum<CS>(acc);
    if (MODE == MODE_FULL) {
        hw = group_sum<CS>(hw);
        kn = group_sum<CS>(kn);
    }
    if (KSD) {
        a1 = group_sum<CS>(a1);
        a2 = group_sum<CS>(a2);
        a3 = group_sum<CS>(a3);
        a4 = group_sum<CS>(a4);
    }
    float ksdp = 0.f;
    if (i < n && cs == 0) {
        float p = acc * (1.0f / (float)n);
        if (MODE == MODE_FULL) p = p / fminf(fmaxf(hw / fmaxf(kn, 1e-6f), 0.1f), 100.0f) * 0.7f;
        const float s = own.y;
        if (KSD) ksdp = s * a1 + 2.0f * rs * (s * a2 - a3) + 2.0f * ib * a4;
        float beta;
        bool last;
        iter_schedule(P, S, g->budget, it, beta, last);
        const float eff = P.base_step * sqrtf(beta);
        const float temp = (P.flags & SVPF_F_SVLD) ? P.temperature : 0.0f;
        const float vn = P.rms_rho * v[i] + (1.0f - P.rms_rho) * p * p;
        float diff = 0.0f;
        if (temp > 1e-6f) {
            float xi;
            if (D.z_pred == nullptr) {
                CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
                xi = rng.normal((uint32_t)i, (uint32_t)S.timestep, SVPF_PURPOSE_SVLD0 + (uint32_t)it);
            } else {
                xi = (it < D.xi_rows) ? D.xi[((size_t)t * D.xi_rows + it) * n + i] : 0.0f;
            }
            diff = sqrtf(2.0f * eff * temp) * xi;
        }
        v[i] = vn;
        h[i] = clamp_logvol(h[i] + eff * p * rsqrtf(vn + P.rms_eps) + diff);
    }
    if (KSD) {
        float* red = reinterpret_cast<float*>(tile);
        __syncthreads();
        float k1[1] = {ksdp};
        block_sum<kGB, 1>(k1, red);
        if (threadIdx.x == 0) part[blockIdx.x] = k1[0];
    }
}

template <int CS>
__global__ __launch_bounds__(kGB) void grid_transport(const svpf_fparams* __restrict__ Pp,
                                                      const svpf_fstate* __restrict__ Sp,
                                                      const svpf_fdesc* __restrict__ Dp,
                                                      const GridScalars* __restrict__ g,
                                                      const float4* __restrict__ rec, float* __restrict__ h,
                                                      float* __restrict__ v, float* __restrict__ part, int it, int t,
                                                      int n) {
    __shared__ __align__(16) float4 tile[kRecTile];
    if (it >= g->budget || g->stop) return;
    const svpf_fparams& P = *Pp;
    const uint32_t F = P.flags;
    float beta;
    bool last;
    iter_schedule(P, *Sp, g->budget, it, beta, last);
    const bool ksd = last || (F & SVPF_F_KSD_EARLY_STOP);
    const int mode = (F & SVPF_F_NEWTON) ? ((F & SVPF_F_FULL_NEWTON) ? MODE_FULL : MODE_NEWTON) : MODE_PLAIN;
    if (mode == MODE_FULL) {
        if (ksd) grid_transport_rows<CS, MODE_FULL, true>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n);
        else grid_transport_rows<CS, MODE_FULL, false>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n);
    } else if (mode == MODE_NEWTON) {
        if (ksd) grid_transport_rows<CS, MODE_NEWTON, true>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n);
        else grid_transport_rows<CS, MODE_NEWTON, false>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n);
    } else {
        if (ksd) grid_transport_rows<CS, MODE_PLAIN, true>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n);
        else grid_transport_rows<CS, MODE_PLAIN, false>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n);
    }
}

__global__ __launch_bounds__(kGB) void grid_ksd(const svpf_fparams* __restrict__ Pp,
                                                const svpf_fstate* __restrict__ Sp, GridScalars* __restrict__ g,
                                                const float* __restrict__ part, int it, int G, int n) {
    __shared__ float red[64];
    if (it >= g->budget || g->stop) return;
    const svpf_fparams& P = *Pp;
    float beta;
    bool last;
    iter_schedule(P, *Sp, g->budget, it, beta, last);
    const bool early = (P.flags & SVPF_F_KSD_EARLY_STOP) != 0;
    float s[1] = {0.f};
    if (last || early)
        for (int b = threadIdx.x; b < G; b += kGB) s[0] += part[b];
    block_sum<kGB, 1>(s, red);
    if (threadIdx.x != 0) return;
    g->total = it + 1;
    if (!(last || early)) return;
    const float inv_n = 1.0f / (float)n;
    const float ksd = sqrtf(fmaxf(s[0] * inv_n * inv_n, 0.0f));
    g->ksd = ksd;
    if (early && g->ksd_last > 0.0f && it + 1 >= P.stein_min) {
        const float rel = (g->ksd_last - ksd) / g->ksd_last;
        if (rel < P.ksd_thr) {
            g->stop = 1;
            return;
        }
    }
    g->ksd_last = ksd;
}

__global__ __launch_bounds__(kGB) void grid_outputs(const svpf_fparams* __restrict__ Pp,
                                                    const svpf_fstate* __restrict__ Sp,
                                                    const svpf_fdesc* __restrict__ Dp, int t,
                                                    const GridScalars* __restrict__ g, float* __restrict__ h,
                                                    const float* __restrict__ lw, float* __restrict__ part, int n) {
    __shared__ float red[64];
    const svpf_fparams& P = *Pp;
    const svpf_fstate& S = *Sp;
    const svpf_fdesc& D = *Dp;
    const uint32_t F = P.flags;
    const int i = blockIdx.x * kGB + threadIdx.x;
    float v[GP_N] = {-3.0e38f, 0.f, 0.f, 0.f, 3.0e38f, -3.0e38f, 0.f, 0.f};
    if (i < n) {
        float hi = h[i];
        // rejuvenation (src/svpf_opt_kernels.cu:1248-1275)
        if ((F & SVPF_F_REJUV) && (F & SVPF_F_GUIDE) && S.timestep > 10 && S.ksd_prev > P.rej_thr) {
            CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
            const float u = D.u_rej ? D.u_rej[(size_t)t * n + i] : rng.uniform((uint32_t)i, (uint32_t)S.timestep, SVPF_PURPOSE_REJUV);
            if (u < P.rej_p) {
                const float z = D.z_rej ? D.z_rej[(size_t)t * n + i] : rng.normal((uint32_t)i, (uint32_t)S.timestep, SVPF_PURPOSE_REJUV);
                hi = clamp_logvol((1.0f - P.rej_blend) * hi + P.rej_blend * (g->guide_mean + g->guide_std * z));
                h[i] = hi;
            }
        }
        v[GP_LWMAX] = lw[i];
        if (D.log_w && D.log_w != lw) D.log_w[i] = lw[i];
        v[GP_LWSUM] = 1.0f;
        v[GP_H] = hi;
        v[GP_VOL] = safe_expf(hi * 0.5f);
        v[GP_HMIN] = hi;
        v[GP_HMAX] = hi;
    }
    block_partials(v, red, part);
}

__global__ __launch_bounds__(kGB) void grid_finish(const svpf_fparams* __restrict__ Pp, svpf_fstate* __restrict__ Sp,
                                                   const svpf_fdesc* __restrict__ Dp, int t,
                                                   const GridScalars* __restrict__ g, const float* __restrict__ part,
                                                   int nb, int n) {
    __shared__ float red[64];
    const svpf_fparams& P = *Pp;
    svpf_fstate& S = *Sp;
    float M = -3.0e38f;
    for (int b = threadIdx.x; b < nb; b += kGB) M = fmaxf(M, part[(size_t)b * GP_N + GP_LWMAX]);
    M = block_max<kGB>(M, red);
    float sums[3] = {0.f, 0.f, 0.f};
    for (int b = threadIdx.x; b < nb; b += kGB) {
        const float* p = part + (size_t)b * GP_N;
        if (p[GP_LWMAX] > -1.0e38f) sums[0] += p[GP_LWSUM] * expf(p[GP_LWMAX] - M);
        sums[1] += p[GP_VOL];
        sums[2] += p[GP_H];
    }
    block_sum<kGB, 3>(sums, red);
    if (threadIdx.x != 0) return;
    const float inv_n = 1.0f / (float)n;
    const float ll = M + logf(fmaxf(sums[0] / (float)n, 1e-10f));
    const float vol = sums[1] * inv_n, hm = sums[2] * inv_n;
    const float ksd = g->ksd;
    const int total = g->total;
    const svpf_fdesc& D = *Dp;
    if (D.out) {
        float* o = D.out + (size_t)t * SVPF_OUTW;
        o[SVPF_O_LOGLIK] = ll;
        o[SVPF_O_VOL] = vol;
        o[SVPF_O_HMEAN] = hm;
        o[SVPF_O_KSD] = ksd;
        o[SVPF_O_BW] = g->bw;
        o[SVPF_O_PRED_LL] = g->pred_ll;
        o[SVPF_O_STEPS] = (float)total;
        o[SVPF_O_BWSQ] = g->bw_sq;
    }
    S.last_loglik = ll;
    S.last_vol = vol;
    S.last_h_mean = hm;
    S.last_bw = g->bw;
    S.last_ksd = ksd;
    S.last_pred_loglik = g->pred_ll;
    S.last_guide_strength = g->gs;
    S.stein_steps_used = total;
    S.h_mean_prev = hm;
    S.vol_prev = vol;
    S.ksd_prev = ksd;
    if ((P.flags & SVPF_F_ADAPT_MU) && S.timestep > 10) {
        const float Pp_ = S.mu_var + P.mu_q;
        const float Rm = P.mu_rscale * g->bw * g->bw;
        const float Kg = Pp_ / (Pp_ + Rm + 1e-8f);
        S.mu_state = fminf(fmaxf(S.mu_state + Kg * (hm - S.mu_state), P.mu_min), P.mu_max);
        S.mu_var = (1.0f - Kg) * Pp_;
    }
    S.y_prev = g->y_t;
    S.timestep += 1;
}

int grid_cs(int n) {
    // lanes per row such that the row grid has about 1024 workgroups (4 per CU)
    int cs = 1;
    while (cs < 64 && (long)n * cs < 1024L * kGB) cs <<= 1;
    return cs;
}

}  // namespace

extern "C" size_t svpf_grid_workspace_floats(int n) {
    const int nb = (n + kGB - 1) / kGB;
    const size_t npad = ((size_t)n + 3) & ~(size_t)3;
    // rec (4n) | mm (npad) | part ((nb + kMaxG) x GP_N) | scalars (32)
    return 4 * (size_t)n + npad + (size_t)(nb + kMaxG) * GP_N + 32;
}

extern "C" size_t svpf_grid_lds_bytes(int n) {
    int ps = 128;
    while (ps < n) ps <<= 1;
    return ((size_t)2 * ps + 4 + kMedianCands) * sizeof(float);
}

// Raises grid_bandwidth's dynamic LDS limit once per process (not a stream
// operation, so it is kept out of the launches a caller may capture).
extern "C" hipError_t svpf_grid_init(void) {
    static hipError_t done = hipErrorNotReady;
    if (done == hipSuccess) return done;
    done = hipFuncSetAttribute((const void*)grid_bandwidth, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)svpf_grid_lds_bytes(kMaxN));
    return done;
}

#define SVPF_GRID_ITER(CSV)                                                                                       \
    case CSV:                                                                                                     \
        grid_gradient<CSV><<<G, kGB, 0, stream>>>(P, S, gsc, h, mm, rec, lw, it, n, npad);                         \
        grid_transport<CSV><<<G, kGB, 0, stream>>>(P, S, D, gsc, rec, h, v, part, it, t, n);                       \
        break;

// Ticks [t0, t1) of one filter on `stream`: per tick 6 + 3 stein_max launches
// (stein_max bounds the device-side Stein budget).  h, hp, v, lw: the filter's
// particle arrays [n]; ws: svpf_grid_workspace_floats(n) floats.
extern "C" hipError_t svpf_launch_grid(const svpf_fparams* P, svpf_fstate* S, const svpf_fdesc* D, float* h,
                                       float* hp, float* v, float* lw, float* ws, int n, int stein_max, int t0,
                                       int t1, hipStream_t stream) {
    if (n < 2 || n > kMaxN || stein_max < 0 || t0 < 0 || t1 < t0) return hipErrorInvalidValue;
    hipError_t e = svpf_grid_init();
    if (e != hipSuccess) return e;
    const int nb = (n + kGB - 1) / kGB;
    const int npad = (n + 3) & ~3;
    const int cs = grid_cs(n);
    const int G = (n + kGB / cs - 1) / (kGB / cs);
    if (G > kMaxG) return hipErrorInvalidValue;
    float4* rec = reinterpret_cast<float4*>(ws);
    float* mm = ws + 4 * (size_t)n;
    float* part = mm + npad;
    GridScalars* gsc = reinterpret_cast<GridScalars*>(part + (size_t)(nb + kMaxG) * GP_N);
    for (int t = t0; t < t1; ++t) {
        grid_prologue<<<1, 64, 0, stream>>>(P, S, D, t, gsc);
        grid_predict<<<nb, kGB, 0, stream>>>(P, S, D, t, gsc, h, hp, lw, part, n);
        grid_bandwidth<<<1, 1024, svpf_grid_lds_bytes(n), stream>>>(P, S, gsc, h, part, nb, n);
        grid_means<<<(npad + kGB - 1) / kGB, kGB, 0, stream>>>(P, gsc, hp, mm, n, npad);
        for (int it = 0; it < stein_max; ++it) {
            switch (cs) {
                SVPF_GRID_ITER(1)
                SVPF_GRID_ITER(2)
                SVPF_GRID_ITER(4)
                SVPF_GRID_ITER(8)
                SVPF_GRID_ITER(16)
                SVPF_GRID_ITER(32)
                SVPF_GRID_ITER(64)
                default: return hipErrorInvalidValue;
            }
            grid_ksd<<<1, kGB, 0, stream>>>(P, S, gsc, part, it, G, n);
        }
        grid_outputs<<<nb, kGB, 0, stream>>>(P, S, D, t, gsc, h, lw, part, n);
        grid_finish<<<1, kGB, 0, stream>>>(P, S, D, t, gsc, part, nb, n);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
