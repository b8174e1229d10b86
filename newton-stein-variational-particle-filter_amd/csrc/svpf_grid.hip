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
//   grid_bandwidth   1 workgroup   predictive loglik; EMA / Silverman bandwidth
//   grid_sort_runs   n/1024        median rule: sorted runs of 1024 particles (LDS)
//   grid_sort_merge  n/256         rank of every particle across the runs -> sorted particles
//   grid_median_round x 2..5  (n-1)/256  median pairwise r^2, one round each: apply the previous round's
//                                  counts to every row's bracket, count pairs at or below 15 thresholds;
//                                  the converging round gathers the candidates
//   grid_median_final 1 x 1024     rounds still needed if any, rank of the candidates -> bandwidth
//   grid_means       n/256         scaled mixture-component means
//   per Stein iteration k < max:
//     grid_gradient  G x 256       the KSD and early-stop decision after iteration k - 1 (every block
//                                  from the same partials); mixture prior (component means streamed
//                                  through LDS tiles) + likelihood + Newton curvature -> transport
//                                  records {h/bw, s, c/bw, curv}
//     grid_transport G x 256       IMQ transport over record tiles in LDS, RMSProp, Langevin noise; KSD
//                                  partials when the iteration needs them
//   grid_outputs     n/256         rejuvenation, output partials
//   grid_finish      1 workgroup   the last iteration's KSD decision, outputs and the host tail
//
// The Stein budget and early stop are decided on the device: iterations past
// them return at once, so a tick is a fixed launch sequence (capturable as one
// graph).  Rows of the O(N^2) kernels are spread CS lanes per row so that G
// covers the CUs; the per-particle math is the engine's (svpf_engine.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>

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

// The KSD / early-stop state available to Stein iteration k: the KSD of the last
// iteration that computed one, the previous KSD of the early-stop rule, the stop
// flag and the iterations run so far.  GridScalars keeps it for k even and odd
// (dec[k & 1]): the gradient launch of iteration k derives dec[k & 1] from
// dec[(k - 1) & 1] and iteration k - 1's KSD partials, so no launch reads a slot
// it writes.
struct KsdState {
    float ksd, ksd_last;
    int stop, total;
};

struct GridScalars {
    float mu, sigma, y_t, y_prev;
    float gs, guide_mean, guide_std, ib;
    float bw, bw_sq, rs, sq;
    float inv_s2, pred_ll;
    int budget, pad;
    KsdState dec[2];
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

// ---------------------------------------------------------------------------
// Median of the pairwise r^2 over the grid.  The same selection as the
// single-workgroup median_pairwise_r2 (svpf_device.hpp): bisection over the
// bit patterns of non-negative floats with row brackets [plo, phi) of the pairs
// still inside [L, H], the first round seeded around the previous tick's value,
// extended downwards / upwards 8x wider while the rank lies outside, then even
// splits; here with 15 thresholds a round, one round per launch over (n-1)/256
// workgroups of one row per lane, and up to kMedCand candidates gathered at the
// end and narrowed by histogram passes (select_candidates, svpf_device.hpp).  Counts are exact integers, so the result is the float32 value of
// rank floor((M-1)/2) among the M = n(n-1)/2 values fl(fl(s_j - s_i)^2).
// ---------------------------------------------------------------------------
constexpr int kMW = 15;            // thresholds per round
constexpr int kMedCand = 8192;     // candidates gathered for the final selection
constexpr int kMedRows = 256;      // rows per round workgroup
constexpr int kRun = 1024;         // sorted run length of the two-kernel sort
constexpr int kMaxMG = (kMaxN + kMedRows - 1) / kMedRows;
constexpr uint32_t kSeedStep = 1u << 19;  // 1/16 octave of r^2 in float bit patterns
enum { MS_SEED = 0, MS_DOWN = 1, MS_UP = 2, MS_SPLIT = 3 };

struct MedState {
    uint32_t L, H, cL, cH;  // bracket [L, H] of bit patterns; pairs below L, pairs at or below H
    int64_t anchor, step;
    int mode, fresh, done, gathered;  // fresh: thresholds not counted yet; gathered: candidates in mcand
};

// round launches per tick for n particles: the rounds a seeded selection needs at that size
// (16 even sub-brackets a round; the seeded first one holds about 1% of the pairs).
// Above 4096 particles one launch more than the rounds need: the launch after the
// last round applies its counts and gathers the candidates over the whole grid,
// which grid_median_final, one workgroup, would otherwise do over every row.
int med_rounds(int n) { return n <= 1024 ? 1 : (n <= 4096 ? 2 : 4); }

// dst[0, cnt) = src[0, cnt) with 16-byte loads, four in flight per thread (src, dst 16-byte aligned)
template <int BS>
SVPF_DEV void lds_fill(float* __restrict__ dst, const float* __restrict__ src, int cnt) {
    const int n4 = cnt >> 2;
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(dst);
    int k = threadIdx.x;
    for (; k + 3 * BS < n4; k += 4 * BS) {
        const float4 a = s4[k], b = s4[k + BS], c = s4[k + 2 * BS], d = s4[k + 3 * BS];
        d4[k] = a;
        d4[k + BS] = b;
        d4[k + 2 * BS] = c;
        d4[k + 3 * BS] = d;
    }
    for (; k < n4; k += BS) d4[k] = s4[k];
    for (int i = (n4 << 2) + threadIdx.x; i < cnt; i += BS) dst[i] = src[i];
}

// initial selection state from the sorted particles s[0, n) and the previous tick's value
SVPF_DEV MedState med_init(const float* s, int n, float est) {
    MedState st;
    const float dmax = s[n - 1] - s[0];
    st.L = 0u;
    st.H = __float_as_uint(dmax * dmax);
    st.cL = 0u;
    st.cH = (uint32_t)(((uint64_t)n * (uint64_t)(n - 1)) / 2);
    if (!(est > 0.0f && est < 3.0e38f)) {
        const float iq = s[(3 * n) / 4] - s[n / 4];
        est = 0.5f * iq * iq;
    }
    const uint32_t eb = (est > 0.0f && est < 3.0e38f) ? __float_as_uint(est) : 0u;
    st.mode = (eb > st.L && eb < st.H) ? MS_SEED : MS_SPLIT;
    st.anchor = (int64_t)eb;
    st.step = (int64_t)kSeedStep;
    st.fresh = 1;
    st.done = (!(st.cH - st.cL > (uint32_t)kMedCand && st.L < st.H)) ? 1 : 0;
    st.gathered = 0;
    return st;
}

SVPF_DEV uint32_t med_need(int n) {
    const uint32_t M = (uint32_t)(((uint64_t)n * (uint64_t)(n - 1)) / 2);
    return (M - 1u) / 2u + 1u;  // count(<= x) must reach rank + 1
}

// thresholds of a state, non-decreasing in q, inside [L, H)
SVPF_DEV uint32_t med_thr(const MedState& st, int q) {
    int64_t x;
    if (st.mode == MS_SEED) x = st.anchor + (int64_t)(q - kMW / 2) * st.step;
    else if (st.mode == MS_DOWN) x = st.anchor - (int64_t)(kMW - q) * st.step;
    else if (st.mode == MS_UP) x = st.anchor + (int64_t)(q + 1) * st.step;
    else return st.L + (uint32_t)(((uint64_t)(st.H - st.L) * (uint64_t)(q + 1)) / (kMW + 1));
    x = x < (int64_t)st.L ? (int64_t)st.L : x;
    x = x > (int64_t)st.H - 1 ? (int64_t)st.H - 1 : x;
    return (uint32_t)x;
}

SVPF_DEV bool med_converged(const MedState& st) { return !(st.cH - st.cL > (uint32_t)kMedCand && st.L < st.H); }

// Applies the totals c[q] (pairs at or below thr(q)) of the state's thresholds:
// the sub-bracket holding the rank, the next mode.  qs / xlo / xhi drive the
// row bracket updates (med_row_update).
SVPF_DEV void med_advance(MedState& st, const uint32_t* c, uint32_t need, int& qs, uint32_t& xlo, uint32_t& xhi) {
    qs = kMW;
    for (int q = kMW - 1; q >= 0; --q)
        if (c[q] >= need) qs = q;
    xlo = qs > 0 ? med_thr(st, qs - 1) : 0u;
    xhi = qs < kMW ? med_thr(st, qs) : 0u;
    const uint32_t clo = qs > 0 ? c[qs - 1] : 0u, chi = qs < kMW ? c[qs] : 0u;
    if ((st.mode == MS_SEED || st.mode == MS_DOWN) && qs == 0) {
        st.anchor = (int64_t)med_thr(st, 0);  // still below the grid: extend it downwards
        st.step *= 8;
        st.mode = MS_DOWN;
    } else if ((st.mode == MS_SEED || st.mode == MS_UP) && qs == kMW) {
        st.anchor = (int64_t)med_thr(st, kMW - 1);
        st.step *= 8;
        st.mode = MS_UP;
    } else {
        st.mode = MS_SPLIT;
    }
    if (qs > 0) {
        st.L = xlo + 1u;
        st.cL = clo;
    }
    if (qs < kMW) {
        st.H = xhi;
        st.cH = chi;
    }
    st.fresh = 0;
    st.done = med_converged(st) ? 1 : 0;
}

SVPF_DEV void med_row_update(const float* s, int i, int qs, uint32_t xlo, uint32_t xhi, int* plo, int* phi) {
    const float si = s[i];
    int a = plo[i], b = phi[i];
    if (qs > 0) a = row_first_gt(s, si, a, b, __uint_as_float(xlo));
    if (qs < kMW) b = row_first_gt(s, si, a, b, __uint_as_float(xhi));
    plo[i] = a;
    phi[i] = b;
}

// counts of row i at or below the thresholds xb (c[q] += pairs (i, j) with fl(fl(s_j - s_i)^2) <= xb[q])
SVPF_DEV void med_row_count(const float* s, int n, int i, const int* plo, const int* phi, const uint32_t (&xb)[kMW],
                            uint32_t (&c)[kMW]) {
    int res[kMW];
    row_search<kMW>(s, n, s[i], plo[i], phi[i], xb, res);
#pragma unroll
    for (int q = 0; q < kMW; ++q) c[q] += (uint32_t)(res[q] - (i + 1));
}

// workgroup totals of c[kMW] into tot[kMW] (LDS), every thread reads them after the call
template <int BS>
SVPF_DEV void med_block_totals(uint32_t (&c)[kMW], uint32_t* tot) {
    if (threadIdx.x < kMW) tot[threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMW; ++q) c[q] = wave_sum_u32(c[q]);
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < kMW; ++q) atomicAdd(&tot[q], c[q]);
    }
    __syncthreads();
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
    // the means are scaled by sqrt(log2(e) / 2 sigma^2) for the Gaussian state prior, by
    // 1 / sqrt(nu_s sigma^2) for the Student-t one
    g->sq = (F & SVPF_F_STUDENT_STATE) ? sqrtf(g->inv_s2 / P.nu_state) : sqrtf(1.4426950408889634f * 0.5f * g->inv_s2);
    g->budget = budget;
    g->dec[0] = KsdState{0.0f, -1.0f, 0, 0};
    g->dec[1] = g->dec[0];
}

// The state available to Stein iteration k (see KsdState): for k > 0 the decision
// after iteration k - 1 from dec[(k - 1) & 1] and that iteration's KSD partials
// kpart[(k - 1) & 1][0, G) (src/svpf_optimized_graph.cu:1261-1290).  Every block of
// a launch computes it the same way (a fixed-order sum), so all take the same
// decision.  Called by every thread of a kGB-thread block.
SVPF_DEV KsdState ksd_state(const svpf_fparams& P, const svpf_fstate& S, const GridScalars* g,
                            const float* __restrict__ kpart, int k, int G, int n, float* red) {
    if (k == 0) return g->dec[0];
    const int j = k - 1;
    const KsdState prev = g->dec[j & 1];
    KsdState st = prev;
    if (j >= g->budget || prev.stop) return st;  // iteration j did not run
    float beta;
    bool last;
    iter_schedule(P, S, g->budget, j, beta, last);
    const bool early = (P.flags & SVPF_F_KSD_EARLY_STOP) != 0;
    st.total = j + 1;
    if (!(last || early)) return st;
    float s[1] = {0.f};
    const float* kp = kpart + (size_t)(j & 1) * kMaxG;
    for (int b = threadIdx.x; b < G; b += kGB) s[0] += kp[b];
    block_sum<kGB, 1>(s, red);
    const float inv_n = 1.0f / (float)n;
    const float ksd = sqrtf(fmaxf(s[0] * inv_n * inv_n, 0.0f));
    st.ksd = ksd;
    if (early && prev.ksd_last > 0.0f && j + 1 >= P.stein_min) {
        const float rel = (prev.ksd_last - ksd) / prev.ksd_last;
        if (rel < P.ksd_thr) {
            st.stop = 1;
            return st;
        }
    }
    st.ksd_last = ksd;
    return st;
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
            if ((F & SVPF_F_STUDENT_STATE) && P.nu_state <= 30.0f) {
                // Student-t state noise (src/svpf_opt_kernels.cu:31-53): z sqrt(k / chi2_k)
                const int k = max((int)(P.nu_state + 0.5f), 3);
                float chi2 = 0.0f;
                for (int c = 0; c < k; ++c) {
                    const float w = rng.normal((uint32_t)i, (uint32_t)S.timestep, SVPF_PURPOSE_STATE_T0 + (uint32_t)c);
                    chi2 = fmaf(w, w, chi2);
                }
                z *= sqrtf((float)k / (chi2 + 1e-8f));
            }
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
                                                       const float* __restrict__ part, int nb, int n) {
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
        bw = bsq = 0.0f;  // grid_median_final writes the bandwidth
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
        if (P.bandwidth_mode == SVPF_BW_MEDIAN) return;
        g->bw = bw;
        g->bw_sq = bsq;
        g->ib = 1.0f / (bw * bw);
        g->rs = 1.0f / bw;
    }
}

// Sort of the predicted particles for the median rule, in two launches.
// grid_sort_runs: workgroup b sorts h[b*1024, b*1024 + 1024) in LDS into runs.
__global__ __launch_bounds__(kRun) void grid_sort_runs(const svpf_fparams* __restrict__ Pp,
                                                       const float* __restrict__ h, float* __restrict__ runs, int n) {
    __shared__ __align__(16) float src[kRun], dst[kRun];
    if (Pp->bandwidth_mode != SVPF_BW_MEDIAN) return;
    const int base = blockIdx.x * kRun, len = min(kRun, n - base);
    if (threadIdx.x < len) src[threadIdx.x] = h[base + threadIdx.x];
    __syncthreads();
    block_sort<kRun>(src, dst, len);
    if (threadIdx.x < len) runs[base + threadIdx.x] = dst[threadIdx.x];
}

// grid_sort_merge: one element per lane; its rank in the merged order is its
// index in its own run plus, in every other run, the elements before it in a
// stable merge (<= x in lower runs, < x in higher runs), so ranks are distinct.
__global__ __launch_bounds__(kGB) void grid_sort_merge(const svpf_fparams* __restrict__ Pp,
                                                       const float* __restrict__ runs, float* __restrict__ srt, int n) {
    extern __shared__ __align__(16) char lds[];
    if (Pp->bandwidth_mode != SVPF_BW_MEDIAN) return;
    float* r = reinterpret_cast<float*>(lds);  // n floats: all runs
    lds_fill<kGB>(r, runs, n);
    __syncthreads();
    const int i = blockIdx.x * kGB + threadIdx.x;
    if (i == 0) srt[n] = 3.402823466e38f;
    if (i >= n) return;
    const int a = i / kRun, nr = (n + kRun - 1) / kRun;
    const float x = r[i];
    int rank = i - a * kRun;
    for (int b = 0; b < nr; ++b) {
        if (b == a) continue;
        const int lo = b * kRun, len = min(kRun, n - lo);
        // elements of run b before x: y <= x (b < a) or y < x (b > a)
        int pos = 0;
        for (int step = kRun; step > 0; step >>= 1) {
            const int q = pos + step;
            if (q <= len) {
                const float y = r[lo + q - 1];
                pos = (b < a ? (y <= x) : (y < x)) ? q : pos;
            }
        }
        rank += pos;
    }
    srt[rank] = x;
}

// tot[q] += sum over the nmg round workgroups of their count q (tot zeroed and
// synchronised by the caller): every load in flight at once, LDS atomics.
template <int BS>
SVPF_DEV void med_sum_partials(const uint32_t* __restrict__ pp, int nmg, uint32_t* tot) {
    for (int k = threadIdx.x; k < nmg * kMW; k += BS) {
        const uint32_t c = pp[k];
        if (c) atomicAdd(&tot[k % kMW], c);
    }
}

// One selection round r of med_rounds(n): workgroup b owns rows [b*256, b*256+256).
// Round 0 builds the initial state and brackets; later rounds first apply the
// previous round's counts.  The round that reaches the final bracket gathers
// its rows' candidates into mcand.
__global__ __launch_bounds__(kMedRows) void grid_median_round(const svpf_fparams* __restrict__ Pp,
                                                              const svpf_fstate* __restrict__ Sp,
                                                              const float* __restrict__ srt, int* __restrict__ plo,
                                                              int* __restrict__ phi, MedState* __restrict__ mstate,
                                                              uint32_t* __restrict__ mpart, float* __restrict__ mcand,
                                                              uint32_t* __restrict__ gcount, int r, int n) {
    extern __shared__ __align__(16) char lds[];
    __shared__ uint32_t tot[kMW];
    __shared__ uint32_t wsum[kMedRows / kWave];
    __shared__ uint32_t cbase;
    if (Pp->bandwidth_mode != SVPF_BW_MEDIAN) return;
    MedState st;
    if (r > 0) {
        st = mstate[(r + 1) & 1];
        if (st.done) return;  // uniform
    }
    float* s = reinterpret_cast<float*>(lds);  // n + 1 floats: sorted particles, +inf at n
    lds_fill<kMedRows>(s, srt, n + 1);
    const int rows = n - 1;
    const int nmg = (rows + kMedRows - 1) / kMedRows;
    const int i = blockIdx.x * kMedRows + threadIdx.x;
    if (r > 0) {
        // totals of the previous round (integer counts: every workgroup gets the same sums)
        if (threadIdx.x < kMW) tot[threadIdx.x] = 0u;
        __syncthreads();
        med_sum_partials<kMedRows>(mpart + (size_t)((r + 1) & 1) * kMaxMG * kMW, nmg, tot);
    }
    __syncthreads();
    if (r == 0) {
        st = med_init(s, n, Sp->med_r2);
        if (i < rows) {
            plo[i] = i + 1;
            phi[i] = n;
        }
        st.fresh = 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *gcount = 0u;
            mstate[1].done = 0;  // no stale result from the previous tick
            mstate[0] = st;
        }
        if (st.done) return;  // uniform: grid_median_final gathers
    } else {
        int qs;
        uint32_t xlo, xhi;
        med_advance(st, tot, med_need(n), qs, xlo, xhi);
        if (i < rows) med_row_update(s, i, qs, xlo, xhi, plo, phi);
        if (st.done) {
            if (st.L < st.H) {
                // gather this workgroup's candidates at an offset reserved from gcount
                const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
                const uint32_t cnt = i < rows ? (uint32_t)(phi[i] - plo[i]) : 0u;
                const uint32_t incl = wave_scan_u32(cnt);
                if (lane == kWave - 1) wsum[w] = incl;
                __syncthreads();
                uint32_t off = incl - cnt, all = 0u;
                for (int q = 0; q < kMedRows / kWave; ++q) {
                    off += (q < w) ? wsum[q] : 0u;
                    all += wsum[q];
                }
                if (threadIdx.x == 0) cbase = all ? atomicAdd(gcount, all) : 0u;
                __syncthreads();
                off += cbase;
                if (i < rows) {
                    const float si = s[i];
                    for (int j = plo[i]; j < phi[i]; ++j) {
                        const float d = s[j] - si;
                        mcand[off++] = d * d;
                    }
                }
                st.gathered = 1;
            }
            if (blockIdx.x == 0 && threadIdx.x == 0) mstate[r & 1] = st;
            return;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) mstate[r & 1] = st;
    }
    uint32_t xb[kMW], c[kMW];
#pragma unroll
    for (int q = 0; q < kMW; ++q) {
        xb[q] = med_thr(st, q);
        c[q] = 0u;
    }
    if (i < rows) med_row_count(s, n, i, plo, phi, xb, c);
    __syncthreads();  // tot was read above
    med_block_totals<kMedRows>(c, tot);
    if (threadIdx.x < kMW) mpart[(size_t)(r & 1) * kMaxMG * kMW + blockIdx.x * kMW + threadIdx.x] = tot[threadIdx.x];
}

// Final selection: rounds still needed (none in the common case) run here over
// all rows; the candidates (gathered by the converging round, or here) are
// ranked by counting.  Writes the bandwidth.
__global__ __launch_bounds__(1024) void grid_median_final(const svpf_fparams* __restrict__ Pp,
                                                          svpf_fstate* __restrict__ Sp, GridScalars* __restrict__ g,
                                                          const float* __restrict__ srt, int* __restrict__ plo,
                                                          int* __restrict__ phi, const MedState* __restrict__ mstate,
                                                          const uint32_t* __restrict__ mpart,
                                                          const float* __restrict__ mcand, int rounds, int n) {
    constexpr int BS = 1024;
    extern __shared__ __align__(16) char lds[];
    __shared__ uint32_t tot[kMW];
    __shared__ uint32_t wsum[BS / kWave];
    __shared__ uint32_t bins[kHistBins];
    __shared__ uint32_t ctr, medv;
    if (Pp->bandwidth_mode != SVPF_BW_MEDIAN) return;
    float* s = reinterpret_cast<float*>(lds);  // n + 1 floats
    float* cand = s + ((n + 4) & ~3);          // kMedCand floats
    const int rows = n - 1;
    const uint32_t need = med_need(n);
    MedState st;
    if (mstate[0].done) st = mstate[0];
    else if (mstate[1].done) st = mstate[1];
    else st = mstate[(rounds - 1) & 1];  // the last round's state, counted but not yet applied
    if (!st.gathered) lds_fill<BS>(s, srt, n + 1);
    __syncthreads();
    if (!st.done) {
        const uint32_t* pp = mpart + (size_t)((rounds - 1) & 1) * kMaxMG * kMW;
        const int nmg = (rows + kMedRows - 1) / kMedRows;
        if (threadIdx.x < kMW) tot[threadIdx.x] = 0u;
        __syncthreads();
        med_sum_partials<BS>(pp, nmg, tot);
        __syncthreads();
        int qs;
        uint32_t xlo, xhi;
        med_advance(st, tot, need, qs, xlo, xhi);
        for (int i = threadIdx.x; i < rows; i += BS) med_row_update(s, i, qs, xlo, xhi, plo, phi);
        __syncthreads();
        while (!st.done) {  // uniform: every thread holds the same state
            uint32_t xb[kMW], c[kMW];
#pragma unroll
            for (int q = 0; q < kMW; ++q) {
                xb[q] = med_thr(st, q);
                c[q] = 0u;
            }
            for (int i = threadIdx.x; i < rows; i += BS) med_row_count(s, n, i, plo, phi, xb, c);
            med_block_totals<BS>(c, tot);
            med_advance(st, tot, need, qs, xlo, xhi);
            for (int i = threadIdx.x; i < rows; i += BS) med_row_update(s, i, qs, xlo, xhi, plo, phi);
            __syncthreads();
        }
    }
    float med;
    if (st.L >= st.H) {
        med = __uint_as_float(st.H);
    } else {
        const int C = (int)(st.cH - st.cL);
        if (st.gathered) {
            for (int k = threadIdx.x; k < C; k += BS) cand[k] = mcand[k];
        } else {
            // gather the candidates (pairs with r^2 in [L, H]) of every row
            const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
            uint32_t cnt = 0u;
            for (int i = threadIdx.x; i < rows; i += BS) cnt += (uint32_t)(phi[i] - plo[i]);
            const uint32_t incl = wave_scan_u32(cnt);
            if (lane == kWave - 1) wsum[w] = incl;
            __syncthreads();
            uint32_t off = incl - cnt;
            for (int q = 0; q < w; ++q) off += wsum[q];
            for (int i = threadIdx.x; i < rows; i += BS) {
                const float si = s[i];
                const int b = phi[i];
                for (int j = plo[i]; j < b; ++j) {
                    const float d = s[j] - si;
                    cand[off++] = d * d;
                }
            }
        }
        __syncthreads();
        med = select_candidates<BS>(cand, C, st.L, st.H, st.cL, need, bins, &ctr, &medv);
    }
    if (threadIdx.x == 0) {
        const float bw = fminf(fmaxf(sqrtf(med), 0.01f), 2.0f);
        Sp->med_r2 = med;
        g->bw = bw;
        g->bw_sq = med;
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

// gradient of rows r = blockIdx.x * (kGB/CS) + threadIdx.x / CS; CS lanes split the components.
// ST: the Student-t state prior (src/svpf_opt_kernels.cu:406-440) in units scaled by
// 1/sqrt(nu_s sigma^2): r_i = exp2(a (log2 q_min - log2 q_i)), q = 1 + dd^2, a = (nu_s + 1)/2.
// save (Heun): the positions at the iteration start are copied there.
template <int CS, bool ST>
__global__ __launch_bounds__(kGB) void grid_gradient(const svpf_fparams* __restrict__ Pp,
                                                     const svpf_fstate* __restrict__ Sp,
                                                     GridScalars* __restrict__ g, const float* __restrict__ h,
                                                     const float* __restrict__ mm, float4* __restrict__ rec,
                                                     float* __restrict__ lw, float* __restrict__ save,
                                                     const float* __restrict__ kpart, int it, int G, int n,
                                                     int npad) {
    __shared__ __align__(16) float tile[kMeanTile];
    __shared__ float kred[64];
    // the early-stop decision after the previous iteration (uniform across the grid)
    const KsdState st = ksd_state(*Pp, *Sp, g, kpart, it, G, n, kred);
    if (it > 0 && blockIdx.x == 0 && threadIdx.x == 0) g->dec[it & 1] = st;
    if (it >= g->budget || st.stop) return;
    const svpf_fparams& P = *Pp;
    const uint32_t F = P.flags;
    const int i = blockIdx.x * (kGB / CS) + threadIdx.x / CS, cs = threadIdx.x % CS;
    float beta;
    bool last;
    iter_schedule(P, *Sp, g->budget, it, beta, last);
    const float sq = g->sq, inv_s2 = g->inv_s2, rs = g->rs;
    const float hi = i < n ? h[i] : 0.f;
    const float hh = hi * sq;
    const float a_t = 0.5f * (P.nu_state + 1.0f);
    // The shift of the log-responsibilities: the reference takes the nearest component's
    // (a first pass over all means).  Any shift at or above it gives the same sums up to
    // rounding as long as no term overflows, and the particle's own component (it was
    // predicted from it) is one: its term is 2^0 and every term is at most 2^own.  So the
    // first pass runs only in blocks with a row whose own component lies more than 10
    // scaled units away (own > 100).
    const float d_own = i < n ? hh - mm[i] : 0.0f;
    const float own = ST ? a_t * __builtin_amdgcn_logf(fmaf(d_own, d_own, 1.0f)) : d_own * d_own;
    const bool exact = __syncthreads_or(own > 100.0f) != 0;
    // pass 1 (blocks with a far row only): nearest component, the reference's max shift;
    // pass 2: the sums.  Each pass streams the means through LDS in tiles.
    float off = exact ? 3.0e38f : own, sr = 0.f, srd = 0.f, shp = 0.f;
    auto sweep = [&](auto&& body) {
        for (int j0 = 0; j0 < npad; j0 += kMeanTile) {
            const int len = min(kMeanTile, npad - j0);
            __syncthreads();
            for (int k = threadIdx.x * 4; k < len; k += kGB * 4)
                *reinterpret_cast<float4*>(tile + k) = *reinterpret_cast<const float4*>(mm + j0 + k);
            __syncthreads();
            const float4* t4 = reinterpret_cast<const float4*>(tile);
            const int nq = len >> 2;
            int q = cs;
            for (; q + CS < nq; q += 2 * CS) {  // two loads in flight
                const float4 m0 = t4[q], m1 = t4[q + CS];
                body(m0);
                body(m1);
            }
            if (q < nq) body(t4[q]);
        }
    };
    if (exact) {
        sweep([&](const float4& m) {
            const float d0 = hh - m.x, d1 = hh - m.y, d2 = hh - m.z, d3 = hh - m.w;
            off = fminf(off, fminf(fminf(d0 * d0, d1 * d1), fminf(d2 * d2, d3 * d3)));
        });
        off = group_min<CS>(off);
        if (ST) off = a_t * __builtin_amdgcn_logf(1.0f + off);
    }
    if (ST) {
        sweep([&](const float4& m) {
            const float dd[4] = {hh - m.x, hh - m.y, hh - m.z, hh - m.w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float q = fmaf(dd[c], dd[c], 1.0f);
                const float iq = __builtin_amdgcn_rcpf(q);
                const float e = __builtin_amdgcn_exp2f(fmaf(-a_t, __builtin_amdgcn_logf(q), off));
                const float u = e * iq;
                sr += e;
                srd = fmaf(u, dd[c], srd);
                shp = fmaf(u * iq, 2.0f - q, shp);
            }
        });
    } else {
        sweep([&](const float4& m) {
            const float d0 = hh - m.x, d1 = hh - m.y, d2 = hh - m.z, d3 = hh - m.w;
            const float e0 = __builtin_amdgcn_exp2f(fmaf(-d0, d0, off));
            const float e1 = __builtin_amdgcn_exp2f(fmaf(-d1, d1, off));
            const float e2 = __builtin_amdgcn_exp2f(fmaf(-d2, d2, off));
            const float e3 = __builtin_amdgcn_exp2f(fmaf(-d3, d3, off));
            sr += (e0 + e1) + (e2 + e3);
            srd = fmaf(e0, d0, fmaf(e1, d1, fmaf(e2, d2, fmaf(e3, d3, srd))));
        });
    }
    sr = group_sum<CS>(sr);
    srd = group_sum<CS>(srd);
    if (ST) shp = group_sum<CS>(shp);
    if (i >= n || cs != 0) return;
    if (save) save[i] = hi;
    const float nu = P.nu, ysq = g->y_t * g->y_t;
    float gp, hpr;
    if (ST) {
        const float w = 1.0f / (sr + 1e-8f), np1s = (P.nu_state + 1.0f) * sq;
        gp = -np1s * srd * w;
        hpr = -np1s * sq * shp * w;
    } else {
        gp = -((srd * (inv_s2 / sq)) / (sr + 1e-8f));
        hpr = -inv_s2;
    }
    const float A = ysq / (safe_expf(hi) + 1e-8f) / nu, opa = 1.0f + A;
    lw[i] = (F & SVPF_F_FAN) ? 0.0f : P.st_const - 0.5f * hi - (nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f));
    const float gl = (F & SVPF_F_EXACT_GRAD) ? -0.5f + 0.5f * (nu + 1.0f) * A / opa - P.lik_offset
                                              : (logf(ysq + 1e-10f) - hi + P.lik_offset) / 1.4f;
    const float eb = (F & SVPF_F_FAN) ? 1.0f : beta;
    const float gg = fminf(fmaxf(gp + eb * gl, -10.0f), 10.0f);
    const float sgn2 = (F & SVPF_F_PAPER_SIGN) ? 2.0f : -2.0f;
    float sj = gg, cj = sgn2, cv = 1.0f;
    if (F & SVPF_F_NEWTON) {
        cv = fminf(fmaxf(0.5f * (nu + 1.0f) * A / (opa * opa) - hpr, 0.1f), 100.0f);
        // the Heun step takes the raw score (src/svpf_optimized_graph.cu:1147-1232)
        if (!(F & SVPF_F_FULL_NEWTON) && !(F & SVPF_F_HEUN)) {
            sj = 0.7f * gg / cv;
            cj = sgn2 / cv;
        }
    }
    rec[i] = make_float4(hi * rs, sj, cj * rs, cv);
}

// Langevin normal of particle i at Stein iteration it (counter generator or pre-drawn row)
SVPF_DEV float grid_xi(const svpf_fdesc& D, const svpf_fstate& S, int i, int it, int t, int n) {
    if (D.z_pred == nullptr) {
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        return rng.normal((uint32_t)i, (uint32_t)S.timestep, SVPF_PURPOSE_SVLD0 + (uint32_t)it);
    }
    return (it < D.xi_rows) ? D.xi[((size_t)t * D.xi_rows + it) * n + i] : 0.0f;
}

// What a transport pass does with phi: PH_UPDATE the RMSProp step; the Heun
// phases (src/svpf_optimized_graph.cu:1147-1232): PH_PRED stores phi1 and moves
// h to the predictor h~ = clamp(h0 + eff phi1 / sqrt(v + e)) (no noise, v kept),
// PH_PHI2 stores (phi1 + phi2) / 2, PH_KSD only the KSD partials.
enum { PH_UPDATE = 0, PH_PRED = 1, PH_PHI2 = 2, PH_KSD = 3 };

template <int CS, int MODE, bool KSD, int PH = PH_UPDATE>
SVPF_DEV void grid_transport_rows(const svpf_fparams& P, const svpf_fstate& S, const svpf_fdesc& D,
                                  const GridScalars* g, const float4* __restrict__ rec, float* __restrict__ h,
                                  float* __restrict__ v, float* __restrict__ part, float4* tile, int it, int t, int n,
                                  const float* __restrict__ ho = nullptr, float* __restrict__ phib = nullptr) {
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
        // In scaled positions the IMQ satisfies d^2 K = 1 - K, so the Stein kernel's factor
        // 4 d^2 K - 1 is 3 - 4K and the full-Newton weight K^2 |3 d^2 - 1| is K |3 - 4K|;
        // on a KSD pass the transport reuses the KSD sums: acc = a1 + sum_j c_j rs d K^2.
        auto body = [&](const float4& e) {
            const float d = ht - e.x;
            const float K = __builtin_amdgcn_rcpf(fmaf(d, d, 1.0f));
            if (KSD) {
                const float K2 = K * K, u = d * K2;
                a1 = fmaf(K, e.y, a1);
                a2 += u;
                a3 = fmaf(e.y, u, a3);
                acc = fmaf(e.z, u, acc);  // repulsion; the attraction a1 joins after the loop
                const float tk = fmaf(-4.0f, K, 3.0f);
                a4 = fmaf(K2, tk, a4);
                if (MODE == MODE_FULL) {
                    hw = fmaf(e.w, K, hw);
                    hw = fmaf(two_ib, K * fabsf(tk), hw);
                    kn += K;
                }
            } else {
                acc = fmaf(K, fmaf(e.z * d, K, e.y), acc);
                if (MODE == MODE_FULL) {
                    hw = fmaf(e.w, K, hw);
                    hw = fmaf(two_ib, K * fabsf(fmaf(-4.0f, K, 3.0f)), hw);
                    kn += K;
                }
            }
        };
        int j = cs;
        for (; j + CS < len; j += 2 * CS) {  // two loads in flight
            const float4 e0 = tile[j], e1 = tile[j + CS];
            body(e0);
            body(e1);
        }
        if (j < len) body(tile[j]);
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
        acc += a1;
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
        if (PH == PH_UPDATE) {
            const float temp = (P.flags & SVPF_F_SVLD) ? P.temperature : 0.0f;
            const float vn = P.rms_rho * v[i] + (1.0f - P.rms_rho) * p * p;
            float diff = 0.0f;
            if (temp > 1e-6f) diff = sqrtf(2.0f * eff * temp) * grid_xi(D, S, i, it, t, n);
            v[i] = vn;
            h[i] = clamp_logvol(h[i] + eff * p * rsqrtf(vn + P.rms_eps) + diff);
        } else if (PH == PH_PRED) {
            phib[i] = p;
            h[i] = clamp_logvol(ho[i] + eff * p * rsqrtf(v[i] + P.rms_eps));
        } else if (PH == PH_PHI2) {
            phib[i] = 0.5f * (phib[i] + p);
        }
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
    if (it >= g->budget || g->dec[it & 1].stop) return;
    part += (size_t)(it & 1) * kMaxG;  // this iteration's KSD partials
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

// The Heun passes of one Stein iteration: PH_PRED and PH_PHI2 with the raw-score
// operator (plain, or full Newton's kernel-weighted curvature); PH_KSD the KSD at
// the original positions with the predicted positions' scores, when this
// iteration needs it.
template <int CS, int PH>
__global__ __launch_bounds__(kGB) void grid_heun_transport(const svpf_fparams* __restrict__ Pp,
                                                           const svpf_fstate* __restrict__ Sp,
                                                           const svpf_fdesc* __restrict__ Dp,
                                                           const GridScalars* __restrict__ g,
                                                           const float4* __restrict__ rec, float* __restrict__ h,
                                                           float* __restrict__ v, float* __restrict__ part,
                                                           const float* __restrict__ ho, float* __restrict__ phib,
                                                           int it, int t, int n) {
    __shared__ __align__(16) float4 tile[kRecTile];
    if (it >= g->budget || g->dec[it & 1].stop) return;
    part += (size_t)(it & 1) * kMaxG;  // this iteration's KSD partials
    const svpf_fparams& P = *Pp;
    const uint32_t F = P.flags;
    float beta;
    bool last;
    iter_schedule(P, *Sp, g->budget, it, beta, last);
    if (PH == PH_KSD) {
        if (!(last || (F & SVPF_F_KSD_EARLY_STOP))) return;
        grid_transport_rows<CS, MODE_PLAIN, true, PH_KSD>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n, ho, phib);
    } else if ((F & SVPF_F_NEWTON) && (F & SVPF_F_FULL_NEWTON)) {
        grid_transport_rows<CS, MODE_FULL, false, PH>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n, ho, phib);
    } else {
        grid_transport_rows<CS, MODE_PLAIN, false, PH>(P, *Sp, *Dp, g, rec, h, v, part, tile, it, t, n, ho, phib);
    }
}

// Heun: the records' positions back at the iteration start (the KSD pass's columns)
__global__ __launch_bounds__(kGB) void grid_heun_ksd_records(const svpf_fparams* __restrict__ Pp,
                                                             const svpf_fstate* __restrict__ Sp,
                                                             const GridScalars* __restrict__ g,
                                                             const float* __restrict__ ho, float4* __restrict__ rec,
                                                             int it, int n) {
    if (it >= g->budget || g->dec[it & 1].stop) return;
    float beta;
    bool last;
    iter_schedule(*Pp, *Sp, g->budget, it, beta, last);
    if (!(last || (Pp->flags & SVPF_F_KSD_EARLY_STOP))) return;
    const int i = blockIdx.x * kGB + threadIdx.x;
    if (i < n) rec[i].x = ho[i] * g->rs;
}

// Heun corrector: from the iteration-start positions with the mean of both operators,
// the RMSProp update and the Langevin noise
__global__ __launch_bounds__(kGB) void grid_heun_correct(const svpf_fparams* __restrict__ Pp,
                                                         const svpf_fstate* __restrict__ Sp,
                                                         const svpf_fdesc* __restrict__ Dp,
                                                         const GridScalars* __restrict__ g,
                                                         const float* __restrict__ ho, const float* __restrict__ phib,
                                                         float* __restrict__ h, float* __restrict__ v, int it, int t,
                                                         int n) {
    if (it >= g->budget || g->dec[it & 1].stop) return;
    const int i = blockIdx.x * kGB + threadIdx.x;
    if (i >= n) return;
    const svpf_fparams& P = *Pp;
    float beta;
    bool last;
    iter_schedule(P, *Sp, g->budget, it, beta, last);
    const float eff = P.base_step * sqrtf(beta);
    const float temp = (P.flags & SVPF_F_SVLD) ? P.temperature : 0.0f;
    const float p = phib[i];
    const float vn = P.rms_rho * v[i] + (1.0f - P.rms_rho) * p * p;
    float diff = 0.0f;
    if (temp > 1e-6f) diff = sqrtf(2.0f * eff * temp) * grid_xi(*Dp, *Sp, i, it, t, n);
    v[i] = vn;
    h[i] = clamp_logvol(ho[i] + eff * p * rsqrtf(vn + P.rms_eps) + diff);
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
                                                   const float* __restrict__ kpart, int stein_max, int G, int nb,
                                                   int n) {
    __shared__ float red[64];
    const svpf_fparams& P = *Pp;
    svpf_fstate& S = *Sp;
    const KsdState fin = ksd_state(P, S, g, kpart, stein_max, G, n, red);
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
    const float ksd = fin.ksd;
    const int total = fin.total;
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


// Workspace of one filter (offsets in floats, each a multiple of 4):
//   rec [n] float4 | mm [npad] | part [(nb + kMaxG) x GP_N] | scalars [32] | srt [n + 4] | runs [n] |
//   plo [n] int | phi [n] int | mstate [2] MedState + gcount (32 floats) | mpart [2][kMaxMG][kMW] uint32 |
//   mcand [kMedCand] | ho [n] | phib [n] | kpart [2][kMaxG]
struct GridWs {
    float4* rec;
    float* mm;
    float* part;
    GridScalars* gsc;
    float* srt;
    float* runs;
    int* plo;
    int* phi;
    MedState* mstate;
    uint32_t* gcount;
    uint32_t* mpart;
    float* mcand;
    float* ho;     // Heun: positions at the iteration start
    float* phib;   // Heun: phi1, then (phi1 + phi2) / 2
    float* kpart;  // [2][kMaxG] KSD block partials of the even / odd Stein iterations
};

size_t grid_ws_layout(int n, float* base, GridWs* w) {
    auto r4 = [](size_t x) { return (x + 3) & ~(size_t)3; };
    const size_t nb = ((size_t)n + kGB - 1) / kGB;
    size_t o = 0;
    const size_t o_rec = o;
    o += 4 * (size_t)n;
    const size_t o_mm = o;
    o += r4(n);
    const size_t o_part = o;
    o += r4((nb + kMaxG) * GP_N);
    const size_t o_gsc = o;
    o += 32;
    const size_t o_srt = o;
    o += r4((size_t)n + 4);
    const size_t o_runs = o;
    o += r4(n);
    const size_t o_plo = o;
    o += r4(n);
    const size_t o_phi = o;
    o += r4(n);
    const size_t o_ms = o;
    o += 32;
    const size_t o_mp = o;
    o += r4(2 * (size_t)kMaxMG * kMW);
    const size_t o_mc = o;
    o += kMedCand;
    const size_t o_ho = o;
    o += r4(n);
    const size_t o_phib = o;
    o += r4(n);
    const size_t o_kpart = o;
    o += 2 * (size_t)kMaxG;
    static_assert(sizeof(GridScalars) <= 32 * sizeof(float) && 2 * sizeof(MedState) <= 28 * sizeof(float), "ws slots");
    if (w) {
        w->rec = reinterpret_cast<float4*>(base + o_rec);
        w->mm = base + o_mm;
        w->part = base + o_part;
        w->gsc = reinterpret_cast<GridScalars*>(base + o_gsc);
        w->srt = base + o_srt;
        w->runs = base + o_runs;
        w->plo = reinterpret_cast<int*>(base + o_plo);
        w->phi = reinterpret_cast<int*>(base + o_phi);
        w->mstate = reinterpret_cast<MedState*>(base + o_ms);
        w->gcount = reinterpret_cast<uint32_t*>(base + o_ms + 28);
        w->mpart = reinterpret_cast<uint32_t*>(base + o_mp);
        w->mcand = base + o_mc;
        w->ho = base + o_ho;
        w->phib = base + o_phib;
        w->kpart = base + o_kpart;
    }
    return o;
}

// dynamic LDS of the merge / round / final kernels
size_t merge_lds_bytes(int n) { return (size_t)n * sizeof(float); }
size_t round_lds_bytes(int n) { return ((size_t)n + 1) * sizeof(float); }
size_t final_lds_bytes(int n) { return ((((size_t)n + 4) & ~(size_t)3) + (size_t)kMedCand) * sizeof(float); }

}  // namespace

extern "C" size_t svpf_grid_workspace_floats(int n) { return grid_ws_layout(n, nullptr, nullptr); }

extern "C" size_t svpf_grid_lds_bytes(int n) {
    return std::max(merge_lds_bytes(n), std::max(round_lds_bytes(n), final_lds_bytes(n)));
}

// Raises the dynamic LDS limits once per process (not a stream operation, so
// it is kept out of the launches a caller may capture).
extern "C" hipError_t svpf_grid_init(void) {
    static hipError_t done = hipErrorNotReady;
    if (done == hipSuccess) return done;
    hipError_t e = hipFuncSetAttribute((const void*)grid_sort_merge, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)merge_lds_bytes(kMaxN));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)grid_median_round, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)round_lds_bytes(kMaxN));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)grid_median_final, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)final_lds_bytes(kMaxN));
    done = e;
    return done;
}

template <int CS, bool ST>
void grid_iteration(const svpf_fparams* P, svpf_fstate* S, const svpf_fdesc* D, const GridWs& W, float* h, float* v,
                    float* lw, bool heun, int G, int nb, int it, int t, int n, int npad, hipStream_t stream) {
    if (!heun) {
        grid_gradient<CS, ST><<<G, kGB, 0, stream>>>(P, S, W.gsc, h, W.mm, W.rec, lw, nullptr, W.kpart, it, G, n,
                                                     npad);
        grid_transport<CS><<<G, kGB, 0, stream>>>(P, S, D, W.gsc, W.rec, h, v, W.kpart, it, t, n);
        return;
    }
    // both gradient launches of a Heun iteration take the same decision (the partials of iteration
    // it - 1 are untouched until the KSD pass of iteration it)
    grid_gradient<CS, ST><<<G, kGB, 0, stream>>>(P, S, W.gsc, h, W.mm, W.rec, lw, W.ho, W.kpart, it, G, n, npad);
    grid_heun_transport<CS, PH_PRED><<<G, kGB, 0, stream>>>(P, S, D, W.gsc, W.rec, h, v, W.kpart, W.ho, W.phib, it, t,
                                                            n);
    grid_gradient<CS, ST><<<G, kGB, 0, stream>>>(P, S, W.gsc, h, W.mm, W.rec, lw, nullptr, W.kpart, it, G, n, npad);
    grid_heun_transport<CS, PH_PHI2><<<G, kGB, 0, stream>>>(P, S, D, W.gsc, W.rec, h, v, W.kpart, W.ho, W.phib, it, t,
                                                            n);
    grid_heun_ksd_records<<<nb, kGB, 0, stream>>>(P, S, W.gsc, W.ho, W.rec, it, n);
    grid_heun_transport<CS, PH_KSD><<<G, kGB, 0, stream>>>(P, S, D, W.gsc, W.rec, h, v, W.kpart, W.ho, W.phib, it, t,
                                                           n);
    grid_heun_correct<<<nb, kGB, 0, stream>>>(P, S, D, W.gsc, W.ho, W.phib, h, v, it, t, n);
}

#define SVPF_GRID_ITER(CSV)                                                                                       \
    case CSV:                                                                                                     \
        if (student) grid_iteration<CSV, true>(P, S, D, W, h, v, lw, heun, G, nb, it, t, n, npad, stream);       \
        else grid_iteration<CSV, false>(P, S, D, W, h, v, lw, heun, G, nb, it, t, n, npad, stream);              \
        break;

// Ticks [t0, t1) of one filter on `stream`: per tick 10 + med_rounds(n) + 2 stein_max launches
// (7 per iteration with the Heun step; stein_max bounds the device-side Stein budget).
// h, hp, v, lw: the filter's particle arrays [n]; ws: svpf_grid_workspace_floats(n) floats;
// ext: SVPF_GRID_EXT_HEUN / SVPF_GRID_EXT_STUDENT as the filter's flags have them.
extern "C" hipError_t svpf_launch_grid(const svpf_fparams* P, svpf_fstate* S, const svpf_fdesc* D, float* h,
                                       float* hp, float* v, float* lw, float* ws, int n, int stein_max, unsigned ext,
                                       int t0, int t1, hipStream_t stream) {
    const bool heun = (ext & SVPF_GRID_EXT_HEUN) != 0, student = (ext & SVPF_GRID_EXT_STUDENT) != 0;
    if (n < 2 || n > kMaxN || stein_max < 0 || t0 < 0 || t1 < t0) return hipErrorInvalidValue;
    hipError_t e = svpf_grid_init();
    if (e != hipSuccess) return e;
    const int nb = (n + kGB - 1) / kGB;
    const int npad = (n + 3) & ~3;
    const int cs = grid_cs(n);
    const int G = (n + kGB / cs - 1) / (kGB / cs);
    const int nmg = (n - 1 + kMedRows - 1) / kMedRows;
    const int rounds = med_rounds(n);
    if (G > kMaxG || nmg > kMaxMG) return hipErrorInvalidValue;
    GridWs W;
    grid_ws_layout(n, ws, &W);
    for (int t = t0; t < t1; ++t) {
        grid_prologue<<<1, 64, 0, stream>>>(P, S, D, t, W.gsc);
        grid_predict<<<nb, kGB, 0, stream>>>(P, S, D, t, W.gsc, h, hp, lw, W.part, n);
        grid_bandwidth<<<1, 1024, 0, stream>>>(P, S, W.gsc, W.part, nb, n);
        grid_sort_runs<<<(n + kRun - 1) / kRun, kRun, 0, stream>>>(P, h, W.runs, n);
        grid_sort_merge<<<nb, kGB, merge_lds_bytes(n), stream>>>(P, W.runs, W.srt, n);
        for (int r = 0; r < rounds; ++r)
            grid_median_round<<<nmg, kMedRows, round_lds_bytes(n), stream>>>(P, S, W.srt, W.plo, W.phi, W.mstate,
                                                                             W.mpart, W.mcand, W.gcount, r, n);
        grid_median_final<<<1, 1024, final_lds_bytes(n), stream>>>(P, S, W.gsc, W.srt, W.plo, W.phi, W.mstate,
                                                                   W.mpart, W.mcand, rounds, n);
        grid_means<<<(npad + kGB - 1) / kGB, kGB, 0, stream>>>(P, W.gsc, hp, W.mm, n, npad);
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
        }
        grid_outputs<<<nb, kGB, 0, stream>>>(P, S, D, t, W.gsc, h, lw, W.part, n);
        grid_finish<<<1, kGB, 0, stream>>>(P, S, D, t, W.gsc, W.part, W.kpart, stein_max, G, nb, n);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
