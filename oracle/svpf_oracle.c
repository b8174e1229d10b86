/*
 * svpf_oracle.c -- scalar FP64 restatement of the SVPF per-tick step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
 * the product.
 *
 * Follows (paths relative to the reference repository root):
 *   src/svpf_optimized_graph.cu:110-302   create defaults, Student-t constants
 *   src/svpf_optimized_graph.cu:915-1389  svpf_step_graph orchestration
 *   src/svpf_opt_kernels.cu               predict / guide / gradient / transport / KSD / outputs
 *   include/svpf_kernels.cuh:509-541      host EKF guide
 *   src/svpf_fully_fused_v2.cu:90-712     the single-block fused step (svpfo_v2_step)
 *
 * Semantics fixed here where the reference is ambiguous:
 *   - Stein transport is a Jacobi update: every phi_i uses the positions of
 *     the iteration's start (the reference reads them from a shared-memory
 *     snapshot, src/svpf_opt_kernels.cu:548-553).
 *   - bandwidth_mode MEDIAN: ell^2 = lower median of the N(N-1)/2 pairwise
 *     squared distances (index floor((M-1)/2) of the ascending order).
 *   - use_linear_beta: beta_k = 0.3 + 0.7 k/(K-1) over K annealing stages.
 *   - use_ksd_early_stop: the KSD is evaluated every iteration and the loop
 *     stops once the relative improvement drops below
 *     ksd_improvement_threshold (the field exists at include/svpf.cuh:509 but
 *     the reference never reads it).
 */
#include "svpf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static double clamp_d(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
/* src/svpf_opt_kernels.cu:20-26 */
static double clamp_logvol(double h) { return clamp_d(h, -15.0, 5.0); }
static double safe_exp(double x) { return exp(x < 20.0 ? x : 20.0); }

/* ------------------------------------------------------------------------ */
/* Counter-based generator (Philox4x32-10, Salmon, Moraes, Dror, Shaw 2011). */
/* ------------------------------------------------------------------------ */
void svpfo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform on (0,1]: 24 random bits, exactly representable in float32. */
double svpfo_u01_from_bits(uint32_t x) { return (double)((x >> 8) + 1u) * (1.0 / 16777216.0); }

/* Box-Muller (cosine branch) from two 32-bit words. */
double svpfo_normal_from_bits(uint32_t a, uint32_t b) {
    double u1 = svpfo_u01_from_bits(a);
    double u2 = (double)(b >> 8) * (1.0 / 16777216.0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Noise purposes (third counter word). */
enum { PURPOSE_PREDICT = 0, PURPOSE_REJUV = 1, PURPOSE_PREDICT_SLOW = 2, PURPOSE_INIT = 255, PURPOSE_INIT_2F = 254,
       PURPOSE_SVLD0 = 16, PURPOSE_SVLD_SLOW0 = 1024, PURPOSE_STATE_T0 = 128, PURPOSE_JOINT_A = 3,
       PURPOSE_JOINT_B = 4, PURPOSE_JOINT_INIT_A = 253, PURPOSE_JOINT_INIT_B = 252 };

static void philox_draw(uint64_t seed, uint32_t filter, uint32_t tick, uint32_t purpose, uint32_t i,
                        uint32_t out[4]) {
    uint32_t ctr[4] = {i, tick, purpose, filter};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    svpfo_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------ */
/* Closed forms                                                              */
/* ------------------------------------------------------------------------ */

/* src/svpf_optimized_graph.cu:123-125 */
/* src/svpf_opt_kernels.cu:31-53: t = z sqrt(nu_int / chi2), chi2 the sum of
 * nu_int squared normals, nu_int = max(round(nu), 3); a normal for nu > 30.
 * z is the predict normal (purpose 0, as the Gaussian noise); normal k of the
 * chi-squared sum is the cosine normal of purpose 128 + k. */
double svpfo_student_t_draw(uint64_t seed, uint32_t filter, uint32_t tick, uint32_t i, double nu_state) {
    uint32_t o[4];
    philox_draw(seed, filter, tick, PURPOSE_PREDICT, i, o);
    double z = svpfo_normal_from_bits(o[0], o[1]);
    if (nu_state > 30.0) return z;
    int nu_int = (int)(nu_state + 0.5);
    if (nu_int < 3) nu_int = 3;
    double chi2 = 0.0;
    for (int k = 0; k < nu_int; ++k) {
        philox_draw(seed, filter, tick, PURPOSE_STATE_T0 + (uint32_t)k, i, o);
        double u = svpfo_normal_from_bits(o[0], o[1]);
        chi2 += u * u;
    }
    return z * sqrt((double)nu_int / (chi2 + 1e-8));
}

double svpfo_student_t_const(double nu) {
    return lgamma((nu + 1.0) / 2.0) - lgamma(nu / 2.0) - 0.5 * log(M_PI * nu);
}

/* src/svpf_optimized_graph.cu:127-138: -(log nu + psi(1/2) - psi_hat(nu/2)),
 * psi_hat the 2-term asymptotic (or the pole form below 1). */
double svpfo_implied_offset(double nu) {
    const double psi_half = -1.9635100260214235;
    double nh = nu / 2.0, psi_nh;
    if (nh >= 1.0) psi_nh = log(nh) - 1.0 / (2.0 * nh) - 1.0 / (12.0 * nh * nh);
    else psi_nh = -0.5772156649 - 1.0 / nh;
    return -(log(nu) + psi_half - psi_nh);
}

/* src/svpf_opt_kernels.cu:474-488: Student-t log weight with the 1e-8 guard. */
double svpfo_log_w(double h, double y, double nu, double st_const) {
    double vol = safe_exp(h);
    double A = (y * y) / (vol + 1e-8) / nu;
    double a = A > -0.999 ? A : -0.999;
    return st_const - 0.5 * h - (nu + 1.0) * 0.5 * log1p(a);
}

/* k-th smallest (0-based) of a[0, m), a reordered: iterative quickselect with
 * a median-of-three pivot and a three-way partition (keys equal to the pivot
 * end the search), O(m) expected.  Same value as a[k] of the sorted array. */
#define SVPFO_SELECT(NAME, T)                                                       \
    static T NAME(T* a, size_t m, size_t k) {                                       \
        size_t lo = 0, hi = m;                                                      \
        while (hi - lo > 1) {                                                       \
            T x = a[lo], y = a[lo + (hi - lo) / 2], z = a[hi - 1];                  \
            T mn = x < y ? x : y, mx = x < y ? y : x;                               \
            T p = mx < z ? mx : z;                                                  \
            p = mn > p ? mn : p;                                                    \
            size_t lt = lo, i = lo, gt = hi;                                        \
            while (i < gt) {                                                        \
                T v = a[i];                                                         \
                if (v < p) { a[i] = a[lt]; a[lt] = v; ++lt; ++i; }                  \
                else if (v > p) { --gt; a[i] = a[gt]; a[gt] = v; }                  \
                else ++i;                                                           \
            }                                                                       \
            if (k < lt) hi = lt;                                                    \
            else if (k >= gt) lo = gt;                                              \
            else return p;                                                          \
        }                                                                           \
        return a[lo];                                                               \
    }
SVPFO_SELECT(select_kth_d, double)
SVPFO_SELECT(select_kth_f, float)

/* Lower median (rank (M-1)/2 of the M = n(n-1)/2 values) of the pairwise squared distances. */
double svpfo_median_r2(const double* h, int n) {
    if (n < 2) return 0.0;
    size_t m = (size_t)n * (size_t)(n - 1) / 2, k = 0;
    double* r = (double*)malloc(m * sizeof(double));
    if (!r) return NAN;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) { double d = h[i] - h[j]; r[k++] = d * d; }
    double med = select_kth_d(r, m, (m - 1) / 2);
    free(r);
    return med;
}

/* Same selection in float32 arithmetic (d = fl(h_i - h_j), r = fl(d*d)); the
 * bit pattern of the result is what the device selection must reproduce. */
float svpfo_median_r2_f32(const float* h, int n) {
    if (n < 2) return 0.0f;
    size_t m = (size_t)n * (size_t)(n - 1) / 2, k = 0;
    float* r = (float*)malloc(m * sizeof(float));
    if (!r) return NAN;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) { volatile float d = h[i] - h[j]; volatile float q = d * d; r[k++] = q; }
    float med = select_kth_f(r, m, (m - 1) / 2);
    free(r);
    return med;
}

/* src/svpf_optimized_graph.cu:1066-1092 */
double svpfo_adaptive_beta(double ksd_prev, int ai, int n_anneal) {
    if (ai == n_anneal - 1) return 1.0;
    const double ksd_high = 0.50, ksd_low = 0.05, beta_min = 0.30, beta_mid = 0.80;
    double beta;
    if (ksd_prev > ksd_high) beta = beta_min;
    else if (ksd_prev < ksd_low) beta = beta_mid;
    else {
        double t = (ksd_high - ksd_prev) / (ksd_high - ksd_low);
        beta = beta_min + t * (beta_mid - beta_min);
    }
    double progress = (double)(ai + 1) / (double)n_anneal;
    beta *= (0.5 + 0.5 * progress);
    return clamp_d(beta, 0.1, 1.0);
}

/* src/svpf_optimized_graph.cu:1097-1112 */
int svpfo_stein_budget(const svpfo_cfg* c, const svpfo_state* s) {
    if (s->timestep < 10) return c->stein_max_steps;
    double x = (s->ksd_prev - 0.05) / (0.50 - 0.05);
    x = clamp_d(x, 0.0, 1.0);
    /* float32 in the reference: (int)(x * (max-min)) truncates toward zero */
    return c->stein_min_steps + (int)((float)x * (float)(c->stein_max_steps - c->stein_min_steps));
}

/* include/svpf_kernels.cuh:509-541 */
void svpfo_ekf_update(const svpfo_cfg* c, svpfo_state* s, double y_t, double mu, double rho, double sigma_z) {
    (void)c;
    if (!s->guide_initialized) {
        s->guide_mean = mu;
        s->guide_var = sigma_z * sigma_z / (1.0 - rho * rho);
        s->guide_initialized = 1;
    }
    double m_pred = mu + rho * (s->guide_mean - mu);
    double P_pred = rho * rho * s->guide_var + sigma_z * sigma_z;
    double log_y2 = log(y_t * y_t + 1e-8);
    double obs_offset = -svpfo_implied_offset(c->nu);
    double R = 4.93 + 2.0;
    double S = P_pred + R;
    double K = P_pred / (S + 1e-8);
    double innov = log_y2 - (m_pred + obs_offset);
    s->guide_mean = m_pred + K * innov;
    s->guide_var = (1.0 - K) * P_pred;
}

/* ------------------------------------------------------------------------ */
/* O(N^2) pieces                                                             */
/* ------------------------------------------------------------------------ */

/* src/svpf_opt_kernels.cu:354-521 (Gaussian state prior branch). */
void svpfo_gradient(const svpfo_cfg* c, const double* h, const double* h_prev, double y_t, double y_prev,
                    double mu, double sigma_z, double beta,
                    double* grad, double* log_w, double* precond, double* curv) {
    int n = c->n;
    double rho = c->rho, gamma = c->gamma, nu = c->nu;
    double st_const = svpfo_student_t_const(nu);
    double* m = (double*)malloc((size_t)n * sizeof(double));
    for (int k = 0; k < n; ++k) {
        double hp = h_prev[k];
        double lev = gamma * y_prev / (exp(hp * 0.5) + 1e-8);
        m[k] = mu + rho * (hp - mu) + lev;
    }
    double s2 = sigma_z * sigma_z;
    double inv_2s2 = 0.5 / s2, inv_s2 = 1.0 / s2;
    for (int j = 0; j < n; ++j) {
        double hj = h[j];
        double g_prior, h_prior;
        if (c->use_student_t_state) {
            /* :406-440: components log r_i = -(nu_s+1)/2 log(1 + d^2/(nu_s sigma^2)),
             * gradient -(nu_s+1) d/(nu_s sigma^2 + d^2), Hessian the r-weighted mean of
             * -(nu_s+1)(nu_s sigma^2 - d^2)/(nu_s sigma^2 + d^2)^2 */
            double ns2 = c->nu_state * s2, np1 = c->nu_state + 1.0;
            double lmax = -1e10;
            for (int i = 0; i < n; ++i) {
                double d = hj - m[i];
                double l = -0.5 * np1 * log(1.0 + d * d / ns2);
                if (l > lmax) lmax = l;
            }
            double sum_r = 0.0, wg = 0.0, wh = 0.0;
            for (int i = 0; i < n; ++i) {
                double d = hj - m[i], den = ns2 + d * d;
                double r = exp(-0.5 * np1 * log(1.0 + d * d / ns2) - lmax);
                sum_r += r;
                wg -= r * np1 * d / den;
                wh += r * (-np1 * (ns2 - d * d) / (den * den));
            }
            g_prior = wg / (sum_r + 1e-8);
            h_prior = wh / (sum_r + 1e-8);
        } else {
            double lmax = -1e10;
            for (int i = 0; i < n; ++i) { double d = hj - m[i]; double l = -d * d * inv_2s2; if (l > lmax) lmax = l; }
            double sum_r = 0.0, wg = 0.0;
            for (int i = 0; i < n; ++i) {
                double d = hj - m[i];
                double r = exp(-d * d * inv_2s2 - lmax);
                sum_r += r;
                wg -= r * d * inv_s2;
            }
            g_prior = wg / (sum_r + 1e-8);
            h_prior = -inv_s2;
        }

        double vol = safe_exp(hj);
        double ysq = y_t * y_t;
        double A = ysq / (vol + 1e-8) / nu;
        double opa = 1.0 + A;
        if (c->use_fan_mode) log_w[j] = 0.0;
        else log_w[j] = svpfo_log_w(hj, y_t, nu, st_const);
        double g_lik;
        if (c->use_exact_gradient) g_lik = -0.5 + 0.5 * (nu + 1.0) * A / opa - c->lik_offset;
        else g_lik = (log(ysq + 1e-10) - hj + c->lik_offset) / 1.4;
        double eb = c->use_fan_mode ? 1.0 : beta;
        double g = clamp_d(g_prior + eb * g_lik, -10.0, 10.0);
        grad[j] = g;
        if (c->use_newton) {
            double h_lik = -0.5 * (nu + 1.0) * A / (opa * opa);
            double cv = clamp_d(-(h_lik + h_prior), 0.1, 100.0);
            curv[j] = cv;
            precond[j] = 0.7 * g / cv;
        } else {
            curv[j] = 1.0;
            precond[j] = g;
        }
    }
    free(m);
}

/* src/svpf_opt_kernels.cu:527-599 (plain), :742-906 (local Newton),
 * :912-1103 (full Newton).  s is the score the transport uses: the raw
 * gradient for plain and full Newton, the preconditioned one for local Newton.
 * curv is used by local Newton (repulsion / curv_j) and full Newton (H-bar). */
void svpfo_stein_phi(const svpfo_cfg* c, const double* h, const double* s, const double* curv,
                     double inv_bw_sq, double* phi, double* ksd_partial) {
    int n = c->n;
    double sgn = (c->stein_sign == 1) ? 1.0 : -1.0;
    int newton = c->use_newton && !c->use_full_newton;
    int full = c->use_newton && c->use_full_newton;
    for (int i = 0; i < n; ++i) {
        double hi = h[i], si = s[i];
        double ks = 0.0, gks = 0.0, ksd = 0.0, Hw = 0.0, Kn = 0.0;
        for (int j = 0; j < n; ++j) {
            double d = hi - h[j];
            double dist = d * d * inv_bw_sq;
            double K = 1.0 / (1.0 + dist);
            double K2 = K * K;
            ks += K * s[j];
            double rep = sgn * 2.0 * d * inv_bw_sq * K2;
            if (newton) rep /= curv[j];
            gks += rep;
            if (full) {
                Hw += curv[j] * K + 2.0 * inv_bw_sq * K2 * fabs(3.0 * dist - 1.0);
                Kn += K;
            }
            double gx = -2.0 * d * inv_bw_sq * K2;
            double gy = -gx;
            double hxy = 2.0 * inv_bw_sq * K2 * (4.0 * dist * K - 1.0);
            ksd += K * si * s[j] + si * gy + s[j] * gx + hxy;
        }
        double p = (ks + gks) / (double)n;
        if (full) {
            double Hb = clamp_d(Hw / (Kn > 1e-6 ? Kn : 1e-6), 0.1, 100.0);
            p = p / Hb * 0.7;
        }
        phi[i] = p;
        if (ksd_partial) ksd_partial[i] = ksd;
    }
}

/* ------------------------------------------------------------------------ */
/* Bandwidth                                                                 */
/* ------------------------------------------------------------------------ */

/* Returns inv_bw_sq; writes bw and ell^2.  Reference EMA rule:
 * src/svpf_opt_kernels.cu:1159-1235.  Silverman: src/svpf_fully_fused_v2.cu:397-424. */
static double svpfo_bandwidth(const svpfo_cfg* c, svpfo_state* st, const double* h, double y_t,
                              double* bw_out, double* raw_out) {
    int n = c->n;
    double bw, raw;
    if (c->bandwidth_mode == SVPFO_BW_MEDIAN) {
        double med = svpfo_median_r2(h, n);
        bw = clamp_d(sqrt(med), 0.01, 2.0);
        raw = med;
    } else if (c->bandwidth_mode == SVPFO_BW_SILVERMAN) {
        double sum = 0.0, sq = 0.0;
        for (int i = 0; i < n; ++i) { sum += h[i]; sq += h[i] * h[i]; }
        double mean = sum / n, var = sq / n - mean * mean;
        double sd = sqrt(var > 1e-6 ? var : 1e-6);
        bw = clamp_d(1.06 * sd * pow((double)n, -0.2), 0.05, 2.0);
        raw = bw * bw;
    } else {
        double sum = 0.0, sq = 0.0, mn = 1e10, mx = -1e10;
        for (int i = 0; i < n; ++i) {
            double v = h[i];
            sum += v; sq += v * v;
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        double mean = sum / n, var = sq / n - mean * mean, spread = mx - mn;
        double bw_sq_new = 2.0 * var / log((double)n + 1.0);
        if (bw_sq_new < 1e-6) bw_sq_new = 1e-6;
        double prev = st->bw_sq_ema;
        double bw_sq = prev > 0.0 ? 0.3 * bw_sq_new + 0.7 * prev : bw_sq_new;
        double ar = fabs(y_t);
        double re = st->return_ema, rv = st->return_var;
        re = re > 0.0 ? 0.05 * ar + 0.95 * re : ar;
        rv = rv > 0.0 ? 0.05 * ar * ar + 0.95 * rv : ar * ar;
        st->return_ema = re;
        st->return_var = rv;
        double vol_ratio = ar / (re > 1e-8 ? re : 1e-8);
        double spread_factor = spread * 0.5 < 2.0 ? spread * 0.5 : 2.0;
        double comb = vol_ratio > spread_factor ? vol_ratio : spread_factor;
        double sc = 1.0 - 0.25 * (comb - 1.0 < 2.0 ? comb - 1.0 : 2.0);
        sc = clamp_d(sc, 0.5, 1.0);
        bw_sq *= sc;
        st->bw_sq_ema = bw_sq;
        bw = clamp_d(sqrt(bw_sq), 0.01, 2.0);
        raw = bw_sq;
    }
    *bw_out = bw;
    *raw_out = raw;
    return 1.0 / (bw * bw);
}

/* ------------------------------------------------------------------------ */
/* Tick                                                                      */
/* ------------------------------------------------------------------------ */

static double noise_normal(const svpfo_noise* nz, const float* arr, int i, uint32_t tick, uint32_t purpose) {
    if (arr) return (double)arr[i];
    uint32_t o[4];
    philox_draw(nz->seed, nz->filter_id, tick, purpose, (uint32_t)i, o);
    return svpfo_normal_from_bits(o[0], o[1]);
}
static double noise_uniform(const svpfo_noise* nz, const float* arr, int i, uint32_t tick, uint32_t purpose) {
    if (arr) return (double)arr[i];
    uint32_t o[4];
    philox_draw(nz->seed, nz->filter_id, tick, purpose, (uint32_t)i, o);
    return svpfo_u01_from_bits(o[2]);
}

static double lse_mean(const double* lw, int n) {
    double mx = -1e10;
    for (int i = 0; i < n; ++i) if (lw[i] > mx) mx = lw[i];
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += exp(lw[i] - mx);
    double m = s / n;
    return mx + log(m > 1e-10 ? m : 1e-10);
}

int svpfo_tick(const svpfo_cfg* c, svpfo_state* st, double* h, double* h_prev, double* v,
               double y_t, double y_prev, const svpfo_noise* nz, svpfo_out* out) {
    int n = c->n;
    uint32_t tick = (uint32_t)st->timestep;
    int rng = (nz->z_pred == NULL);
    double* grad = (double*)calloc((size_t)n, sizeof(double));
    double* log_w = (double*)calloc((size_t)n, sizeof(double));
    double* prec = (double*)calloc((size_t)n, sizeof(double));
    double* curv = (double*)calloc((size_t)n, sizeof(double));
    double* phi = (double*)calloc((size_t)n, sizeof(double));
    double* kp = (double*)calloc((size_t)n, sizeof(double));
    double* hn = (double*)calloc((size_t)n, sizeof(double));
    if (!grad || !log_w || !prec || !curv || !phi || !kp || !hn) return -1;

    /* 1. effective mu / sigma (src/svpf_optimized_graph.cu:932-948) */
    double mu = c->use_adaptive_mu ? st->mu_state : c->mu;
    double sigma = c->sigma_z;
    if (c->use_adaptive_sigma && st->timestep > 0) {
        double vol_est = st->vol_prev > 1e-4 ? st->vol_prev : 1e-4;
        double rz = fabs(y_t) / vol_est;
        double boost = 1.0;
        if (rz > c->sigma_boost_threshold) {
            double sev = (rz - c->sigma_boost_threshold) / 3.0;
            if (sev > 1.0) sev = 1.0;
            boost = 1.0 + (c->sigma_boost_max - 1.0) * sev;
        }
        sigma = c->sigma_z * boost;
        st->sigma_z_effective = sigma;
    }
    double rho_up = c->use_asymmetric_rho ? c->rho_up : c->rho;
    double rho_down = c->use_asymmetric_rho ? c->rho_down : c->rho;
    double d_rho = c->use_local_params ? c->delta_rho : 0.0;
    double d_sig = c->use_local_params ? c->delta_sigma : 0.0;
    double implied = svpfo_implied_offset(c->nu);

    /* 2. predict (src/svpf_opt_kernels.cu:125-276) */
    for (int i = 0; i < n; ++i) {
        double hi = h[i], hpi = h_prev[i];
        double z = (rng && c->use_student_t_state)
                       ? svpfo_student_t_draw(nz->seed, nz->filter_id, tick, (uint32_t)i, c->nu_state)
                       : noise_normal(nz, nz->z_pred, i, tick, PURPOSE_PREDICT);
        double lev = c->gamma * y_prev / (safe_exp(hi / 2.0) + 1e-8);
        double hnew;
        if (c->use_guided || c->use_mim) {
            double u = rng ? noise_uniform(nz, NULL, i, tick, PURPOSE_PREDICT)
                           : (nz->u_pred ? (double)nz->u_pred[i] : 1.0);
            double dev = hi - st->h_mean_prev;
            double rho_adj = d_rho * tanh(dev);
            double sig_scale = 1.0 + d_sig * fabs(dev);
            double scale = (u < c->mim_jump_prob) ? c->mim_jump_scale : 1.0;
            double base_rho = (hi > hpi) ? rho_up : rho_down;
            double r = clamp_d(base_rho + rho_adj, 0.0, 0.999);
            double sl = sigma * sig_scale;
            double mean_prior = mu + r * (hi - mu) + lev;
            if (c->use_guided) {
                double log_y2 = log(y_t * y_t + 1e-10);
                double mi = log_y2 + implied;
                if (mi < -5.0) mi = -5.0;
                double zs = (mi - mean_prior) / 2.5;
                double act = zs > c->guided_innovation_threshold ? tanh(zs - c->guided_innovation_threshold) : 0.0;
                double a = c->guided_alpha_base + (c->guided_alpha_shock - c->guided_alpha_base) * act;
                double mp = (1.0 - a) * mean_prior + a * mi;
                hnew = clamp_logvol(mp + sl * scale * z);
            } else {
                hnew = clamp_logvol(mean_prior + sl * scale * z);
            }
        } else {
            hnew = clamp_logvol(mu + c->rho * (hi - mu) + sigma * z + lev);
        }
        h_prev[i] = hi;
        h[i] = hnew;
    }

    /* predictive loglik before any transport (loss of the differentiable form) */
    {
        double stc = svpfo_student_t_const(c->nu);
        for (int i = 0; i < n; ++i) log_w[i] = svpfo_log_w(h[i], y_t, c->nu, stc);
        out->pred_loglik = lse_mean(log_w, n);
    }

    /* 3. guide (src/svpf_optimized_graph.cu:1006-1039) */
    if (c->use_guide) {
        double gs = c->guide_strength_base;
        if (c->use_adaptive_guide && st->timestep > 0) {
            double vol_est = st->vol_prev > 1e-4 ? st->vol_prev : 1e-4;
            double rz = fabs(y_t) / vol_est;
            double ih = log(y_t * y_t + 1e-8) + implied;
            double he = log(vol_est * vol_est + 1e-8);
            if (ih - he > 0.0 && rz > c->guide_innovation_threshold) {
                double sev = (rz - c->guide_innovation_threshold) / 3.0;
                if (sev > 1.0) sev = 1.0;
                gs = c->guide_strength_base + (c->guide_strength_max - c->guide_strength_base) * sev;
            }
        }
        svpfo_ekf_update(c, st, y_t, mu, c->rho, c->sigma_z);
        for (int i = 0; i < n; ++i) {
            if (c->use_guide_preserving) {
                double cm = st->h_mean_prev;
                double nm = (1.0 - gs) * cm + gs * st->guide_mean;
                h[i] = clamp_logvol(nm + (h[i] - cm));
            } else {
                h[i] = clamp_logvol(h[i] + gs * (st->guide_mean - h[i]));
            }
        }
    }

    /* 4. bandwidth */
    double bw, bw_raw;
    double inv_bw_sq = svpfo_bandwidth(c, st, h, y_t, &bw, &bw_raw);

    /* 5. annealed Stein loop (src/svpf_optimized_graph.cu:1057-1317) */
    int n_anneal = c->use_annealing ? c->n_anneal_steps : 1;
    if (n_anneal < 1) n_anneal = 1;
    double base_step = 0.1 * (c->use_guide ? 0.5 : 1.0);
    int budget = svpfo_stein_budget(c, st);
    int total = 0, stop = 0;
    double ksd = 0.0, ksd_last_iter = -1.0;
    int use_newton = c->use_newton, full = c->use_newton && c->use_full_newton;
    for (int ai = 0; ai < n_anneal && !stop; ++ai) {
        double beta;
        if (!c->use_annealing) beta = 1.0;
        else if (c->use_linear_beta) beta = n_anneal > 1 ? 0.3 + 0.7 * (double)ai / (double)(n_anneal - 1) : 1.0;
        else if (c->use_adaptive_beta) beta = svpfo_adaptive_beta(st->ksd_prev, ai, n_anneal);
        else { static const double sch[3] = {0.3, 0.65, 1.0}; beta = sch[ai % 3]; }
        double bf = sqrt(beta);
        double temp = c->use_svld ? c->temperature : 0.0;
        int si = budget / n_anneal;
        if (ai == n_anneal - 1) si = budget - si * (n_anneal - 1);
        for (int s = 0; s < si; ++s) {
            int last = (ai == n_anneal - 1) && (s == si - 1);
            int row = total;
            total++;
            int want_ksd = last || c->use_ksd_early_stop;
            double eff = base_step * bf;
            if (c->use_heun) {
                /* src/svpf_optimized_graph.cu:1147-1232, src/svpf_opt_kernels.cu:1293-1566.  The
                 * Stein operator takes the raw score (plain form, or the kernel-weighted Hessian of
                 * full Newton); predictor: Euler step from h with the current RMSProp state and no
                 * noise; corrector: from h with the mean of the two operators, the RMSProp update
                 * and the Langevin noise; KSD at the original positions with the score of the
                 * predicted ones. */
                svpfo_cfg ch = *c;
                ch.use_newton = full;
                ch.use_full_newton = full;
                svpfo_cfg cp = *c;
                cp.use_newton = 0;
                cp.use_full_newton = 0;
                double* ho = (double*)malloc((size_t)n * sizeof(double));
                double* phi2 = (double*)malloc((size_t)n * sizeof(double));
                memcpy(ho, h, (size_t)n * sizeof(double));
                svpfo_gradient(c, h, h_prev, y_t, y_prev, mu, sigma, beta, grad, log_w, prec, curv);
                svpfo_stein_phi(&ch, h, grad, curv, inv_bw_sq, phi, NULL);
                for (int i = 0; i < n; ++i) h[i] = clamp_logvol(ho[i] + eff * phi[i] / sqrt(v[i] + c->rmsprop_eps));
                svpfo_gradient(c, h, h_prev, y_t, y_prev, mu, sigma, beta, grad, log_w, prec, curv);
                svpfo_stein_phi(&ch, h, grad, curv, inv_bw_sq, phi2, NULL);
                if (want_ksd) svpfo_stein_phi(&cp, ho, grad, curv, inv_bw_sq, hn, kp);
                for (int i = 0; i < n; ++i) {
                    double p = 0.5 * (phi[i] + phi2[i]);
                    double vn = c->rmsprop_rho * v[i] + (1.0 - c->rmsprop_rho) * p * p;
                    v[i] = vn;
                    double drift = eff * p / sqrt(vn + c->rmsprop_eps);
                    double diff = 0.0;
                    if (temp > 1e-6) {
                        double xi;
                        if (rng) xi = noise_normal(nz, NULL, i, tick, PURPOSE_SVLD0 + (uint32_t)row);
                        else xi = (row < nz->xi_rows) ? (double)nz->xi[(size_t)row * n + i] : 0.0;
                        diff = sqrt(2.0 * eff * temp) * xi;
                    }
                    hn[i] = clamp_logvol(ho[i] + drift + diff);
                }
                free(ho);
                free(phi2);
            } else {
            svpfo_gradient(c, h, h_prev, y_t, y_prev, mu, sigma, beta, grad, log_w, prec, curv);
            const double* sc = (use_newton && !full) ? prec : grad;
            svpfo_stein_phi(c, h, sc, curv, inv_bw_sq, phi, want_ksd ? kp : NULL);
            for (int i = 0; i < n; ++i) {
                double p = phi[i];
                double vn = c->rmsprop_rho * v[i] + (1.0 - c->rmsprop_rho) * p * p;
                v[i] = vn;
                double drift = eff * p / sqrt(vn + c->rmsprop_eps);
                double diff = 0.0;
                if (temp > 1e-6) {
                    double xi;
                    if (rng) xi = noise_normal(nz, NULL, i, tick, PURPOSE_SVLD0 + (uint32_t)row);
                    else xi = (row < nz->xi_rows) ? (double)nz->xi[(size_t)row * n + i] : 0.0;
                    diff = sqrt(2.0 * eff * temp) * xi;
                }
                hn[i] = clamp_logvol(h[i] + drift + diff);
            }
            }
            memcpy(h, hn, (size_t)n * sizeof(double));
            if (want_ksd) {
                double sum = 0.0;
                for (int i = 0; i < n; ++i) sum += kp[i];
                double k2 = sum / ((double)n * (double)n);
                ksd = sqrt(k2 > 0.0 ? k2 : 0.0);
                if (c->use_ksd_early_stop && ksd_last_iter > 0.0 && total >= c->stein_min_steps) {
                    double rel = (ksd_last_iter - ksd) / ksd_last_iter;
                    if (rel < c->ksd_improvement_threshold) { stop = 1; break; }
                }
                ksd_last_iter = ksd;
            }
        }
    }
    st->stein_steps_used = total;

    /* 6. rejuvenation (src/svpf_optimized_graph.cu:1328-1341, src/svpf_opt_kernels.cu:1248-1275) */
    if (c->use_rejuvenation && c->use_guide && st->timestep > 10 && st->ksd_prev > c->rejuv_ksd_threshold) {
        double gsd = sqrt(st->guide_var > 1e-6 ? st->guide_var : 1e-6);
        for (int i = 0; i < n; ++i) {
            double u = noise_uniform(nz, nz->u_rej, i, tick, PURPOSE_REJUV);
            if (u < c->rejuv_prob) {
                double z = noise_normal(nz, nz->z_rej, i, tick, PURPOSE_REJUV);
                h[i] = clamp_logvol((1.0 - c->rejuv_blend) * h[i] + c->rejuv_blend * (st->guide_mean + gsd * z));
            }
        }
    }

    /* 7. outputs (src/svpf_opt_kernels.cu:1109-1153) */
    double sv = 0.0, sh = 0.0;
    for (int i = 0; i < n; ++i) { sv += safe_exp(h[i] * 0.5); sh += h[i]; }
    out->loglik = lse_mean(log_w, n);
    out->vol = sv / n;
    out->h_mean = sh / n;
    out->ksd = ksd;
    out->bandwidth = bw;
    out->bw_sq = bw_raw;
    out->steps = total;

    /* 8. host tail (src/svpf_optimized_graph.cu:1383-1388) */
    st->h_mean_prev = out->h_mean;
    st->vol_prev = out->vol;
    st->ksd_prev = ksd;
    if (c->use_adaptive_mu && st->timestep > 10) {
        double Pp = st->mu_var + c->mu_process_var;
        double R = c->mu_obs_var_scale * bw * bw;
        double K = Pp / (Pp + R + 1e-8);
        double mn = st->mu_state + K * (out->h_mean - st->mu_state);
        st->mu_state = clamp_d(mn, c->mu_min, c->mu_max);
        st->mu_var = (1.0 - K) * Pp;
    }
    st->timestep++;

    free(grad); free(log_w); free(prec); free(curv); free(phi); free(kp); free(hn);
    return 0;
}

int svpfo_run(const svpfo_cfg* c, svpfo_state* st, double* h, double* h_prev, double* v,
              int T, const float* y, float y_prev0,
              const float* z_pred, const float* u_pred, const float* xi, int xi_rows,
              const float* u_rej, const float* z_rej,
              double* loglik, double* pred_loglik, double* vol, double* h_mean,
              double* ksd, double* bw, int* steps) {
    int n = c->n;
    double yp = y_prev0;
    for (int t = 0; t < T; ++t) {
        svpfo_noise nz;
        memset(&nz, 0, sizeof(nz));
        nz.z_pred = z_pred ? z_pred + (size_t)t * n : NULL;
        nz.u_pred = u_pred ? u_pred + (size_t)t * n : NULL;
        nz.xi = xi ? xi + (size_t)t * xi_rows * n : NULL;
        nz.xi_rows = xi_rows;
        nz.u_rej = u_rej ? u_rej + (size_t)t * n : NULL;
        nz.z_rej = z_rej ? z_rej + (size_t)t * n : NULL;
        svpfo_out o;
        int rc = svpfo_tick(c, st, h, h_prev, v, (double)y[t], yp, &nz, &o);
        if (rc) return rc;
        if (loglik) loglik[t] = o.loglik;
        if (pred_loglik) pred_loglik[t] = o.pred_loglik;
        if (vol) vol[t] = o.vol;
        if (h_mean) h_mean[t] = o.h_mean;
        if (ksd) ksd[t] = o.ksd;
        if (bw) bw[t] = o.bandwidth;
        if (steps) steps[t] = o.steps;
        yp = (double)y[t];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The single-block fused step, src/svpf_fully_fused_v2.cu:90-712            */
/* ------------------------------------------------------------------------ */
/* Its semantics differ from svpf_step_graph's: the EKF guide uses the
 * adaptive mu and re-initialises when guide_var < 1e-6 (:180-263); local
 * parameters apply whenever delta_rho or delta_sigma is positive and MIM
 * whenever the jump draw falls under mim_jump_prob (:282-327); the bandwidth is
 * Silverman's on a two-pass variance (:397-424); beta ramps 0.3 + 0.5 ai/(A-1)
 * with the last stage at 1 and the remainder iterations in the last stage
 * (:430-455); the prior gradient is clamped to +-10 on its own (:466-484); full
 * Newton is selected by use_full_newton alone, and use_newton scales the
 * repulsion by the stored per-particle Hessian term (:534-574); the KSD of the
 * last iteration takes raw scores, undoing the Newton preconditioning
 * (:579-596); rejuvenation does not need the guide (:640-652); adaptive mu is
 * an EMA clamped to [-10, 0] (:700-706). */
static double v2_normal(const svpfo_noise* nz, const float* arr, int i, uint32_t tick, uint32_t purpose,
                        double missing) {
    if (nz->z_pred == NULL) return noise_normal(nz, NULL, i, tick, purpose);
    return arr ? (double)arr[i] : missing;
}
static double v2_uniform(const svpfo_noise* nz, const float* arr, int i, uint32_t tick, uint32_t purpose,
                         double missing) {
    if (nz->z_pred == NULL) return noise_uniform(nz, NULL, i, tick, purpose);
    return arr ? (double)arr[i] : missing;
}

int svpfo_v2_step(const svpfo_v2_args* a, double* h, double* h_prev, double* grad, double* log_w, double* v,
                  double* inv_hess, const svpfo_noise* nz, svpfo_v2_io* io) {
    const int n = a->n;
    if (n < 1) return 1;
    const uint32_t tick = (uint32_t)a->timestep;
    const double rho = a->rho, nu = a->nu;

    /* phase 0: scalars (:180-263) */
    const double mu_eff = a->use_adaptive_mu ? io->mu_state : a->mu;
    double gm_prev = io->guide_mean, gvar = io->guide_var;
    if (gvar < 1e-6) {
        gm_prev = a->mu;
        gvar = a->sigma_z * a->sigma_z / (1.0 - rho * rho);
    }
    double boost = 1.0;
    if (a->use_adaptive_sigma && a->timestep > 0) {
        double rz = fabs(a->y_t) / (a->vol_prev > 1e-4 ? a->vol_prev : 1e-4);
        if (rz > a->sigma_boost_threshold) {
            double sev = (rz - a->sigma_boost_threshold) / 3.0;
            boost = 1.0 + (a->sigma_boost_max - 1.0) * (sev < 1.0 ? sev : 1.0);
        }
    }
    const double sig = a->sigma_z * boost;
    double gmean, gstr;
    if (a->use_guide) {
        double m_pred = mu_eff + rho * (gm_prev - mu_eff);
        double P_pred = rho * rho * gvar + a->sigma_z * a->sigma_z;
        double Sv = P_pred + (4.93 + 2.0);
        double K = P_pred / (Sv + 1e-8);
        double innov = log(a->y_t * a->y_t + 1e-8) - (m_pred - a->student_t_implied_offset);
        gmean = m_pred + K * innov;
        gvar = (1.0 - K) * P_pred;
        io->guide_mean = gmean;
        io->guide_var = gvar;
        double rz = fabs(a->y_t) / (a->vol_prev > 1e-4 ? a->vol_prev : 1e-4);
        gstr = a->guide_strength_base;
        if (rz > a->guide_innovation_threshold) {
            double sev = (rz - a->guide_innovation_threshold) / 3.0;
            gstr += (a->guide_strength_max - a->guide_strength_base) * (sev < 1.0 ? sev : 1.0);
        }
    } else {
        gmean = a->h_mean_prev;
        gvar = 1.0;
        gstr = 0.0;
    }

    /* phase 1: mixture means from the previous posterior, predict (:270-370) */
    double* m = (double*)malloc((size_t)n * sizeof(double));
    double* s = (double*)malloc((size_t)n * sizeof(double));
    double* hs = (double*)malloc((size_t)n * sizeof(double));
    double* hn = (double*)malloc((size_t)n * sizeof(double));
    double* kp = (double*)calloc((size_t)n, sizeof(double));
    for (int k = 0; k < n; ++k)
        m[k] = mu_eff + rho * (h[k] - mu_eff) + a->gamma * a->y_prev / (exp(h[k] * 0.5) + 1e-8);
    const int local = a->delta_rho > 0.0 || a->delta_sigma > 0.0;
    for (int i = 0; i < n; ++i) {
        double ho = h[i], lr = rho, ls = sig;
        if (local) {
            double dev = ho - a->h_mean_prev;
            lr = clamp_d(rho + a->delta_rho * tanh(dev), 0.0, 0.999);
            ls = sig * (1.0 + a->delta_sigma * fabs(dev));
        }
        double lev = fabs(a->gamma) > 1e-6 ? a->gamma * a->y_prev / (exp(ho * 0.5) + 1e-8) : 0.0;
        double pm = mu_eff + lr * (ho - mu_eff) + lev;
        double z = v2_normal(nz, nz->z_pred, i, tick, PURPOSE_PREDICT, 0.0);
        double u = v2_uniform(nz, nz->u_pred, i, tick, PURPOSE_PREDICT, 1.0);
        double ms = ls * (1.0 + (u < a->mim_jump_prob ? 1.0 : 0.0) * (a->mim_jump_scale - 1.0));
        if (a->use_guided_predict) {
            double mi = log(a->y_t * a->y_t + 1e-10) + a->student_t_implied_offset;
            if (mi < -5.0) mi = -5.0;
            double zs = (mi - pm) / 2.5;
            double act = zs > a->guided_innovation_thresh_predict ? tanh(zs - a->guided_innovation_thresh_predict) : 0.0;
            double al = a->guided_alpha_base + (a->guided_alpha_shock - a->guided_alpha_base) * act;
            pm = (1.0 - al) * pm + al * mi;
        }
        h_prev[i] = ho;
        h[i] = clamp_logvol(pm + ms * z);
    }

    /* phase 2: guide (:372-390) */
    if (a->use_guide && gstr > 0.0) {
        for (int i = 0; i < n; ++i) {
            double x = a->use_guide_preserving ? h[i] + gstr * (gmean - a->h_mean_prev)
                                               : (1.0 - gstr) * h[i] + gstr * gmean;
            h[i] = clamp_logvol(x);
        }
    }

    /* phase 3: Silverman bandwidth on the two-pass variance (:397-424) */
    double mean = 0.0, var = 0.0;
    for (int i = 0; i < n; ++i) mean += h[i];
    mean /= n;
    for (int i = 0; i < n; ++i) var += (h[i] - mean) * (h[i] - mean);
    var /= n;
    double bw = clamp_d(1.06 * sqrt(var > 1e-6 ? var : 1e-6) * pow((double)n, -0.2), 0.05, 2.0);
    io->bandwidth = bw;
    const double ib = 1.0 / (bw * bw);

    /* phase 4: Stein loop (:430-630) */
    const int na = a->n_anneal_steps < 1 ? 1 : a->n_anneal_steps;
    const int ns = a->n_stein_steps < 1 ? 1 : a->n_stein_steps;
    const int spa = ns / na, rem = ns - spa * na;
    const double sgn = a->stein_sign_mode == 1 ? 1.0 : -1.0;
    const double inv_s2 = 1.0 / (sig * sig), i2s2 = 0.5 * inv_s2;
    int step = 0;
    for (int ai = 0; ai < na; ++ai) {
        double beta = (na <= 1 || ai == na - 1) ? 1.0 : 0.3 + 0.5 * (double)ai / (double)(na - 1);
        int si = spa + (ai == na - 1 ? rem : 0);
        for (int it = 0; it < si; ++it) {
            ++step;
            const int last = step == ns;
            for (int j = 0; j < n; ++j) {
                double hj = h[j], lmax = -1e10;
                for (int k = 0; k < n; ++k) {
                    double d = hj - m[k], l = -d * d * i2s2;
                    if (l > lmax) lmax = l;
                }
                double sr = 0.0, wg = 0.0;
                for (int k = 0; k < n; ++k) {
                    double d = hj - m[k], r = exp(-d * d * i2s2 - lmax);
                    sr += r;
                    wg -= r * d * inv_s2;
                }
                double gp = clamp_d(wg / (sr + 1e-8), -10.0, 10.0);
                double vol = exp(hj * 0.5);
                double A = a->y_t * a->y_t / (vol * vol + 1e-8) / nu, opa = 1.0 + A;
                log_w[j] = a->student_t_const - 0.5 * hj - 0.5 * (nu + 1.0) * log(opa);
                double gl = -0.5 + 0.5 * (nu + 1.0) * A / opa - a->lik_offset;
                double g = clamp_d(gp + beta * gl, -10.0, 10.0);
                grad[j] = g;
                double cv = clamp_d(-(-0.5 * (nu + 1.0) * A / (opa * opa) - inv_s2), 0.1, 100.0);
                if (a->use_newton || a->use_full_newton) inv_hess[j] = cv;
                if (a->use_full_newton) {
                    s[j] = g;
                    hs[j] = cv;
                } else if (a->use_newton) {
                    s[j] = 0.7 * g / cv;
                    hs[j] = 1.0 / cv;
                } else {
                    s[j] = g;
                    hs[j] = 1.0;
                }
            }
            const double eff = a->step_size * sqrt(beta);
            for (int i = 0; i < n; ++i) {
                double ks = 0.0, gk = 0.0, hw = 0.0, kn = 0.0, kq = 0.0;
                for (int j = 0; j < n; ++j) {
                    double d = h[i] - h[j], dist = d * d * ib, K = 1.0 / (1.0 + dist), K2 = K * K;
                    ks += K * s[j];
                    gk += sgn * 2.0 * d * ib * K2 * (a->use_newton ? hs[j] : 1.0);
                    if (a->use_full_newton) {
                        hw += hs[j] * K + 2.0 * ib * K2 * fabs(3.0 * dist - 1.0);
                        kn += K;
                    }
                    if (last) {
                        double ri = grad[i], rj = a->use_newton ? s[j] / (0.7 * hs[j] + 1e-8) : s[j];
                        double gx = -2.0 * d * ib * K2;
                        kq += K * ri * rj - ri * gx + rj * gx + 2.0 * ib * K2 * (4.0 * dist * K - 1.0);
                    }
                }
                double phi = (ks + gk) / n;
                if (a->use_full_newton) phi = phi / clamp_d(hw / (kn > 1e-6 ? kn : 1e-6), 0.1, 100.0) * 0.7;
                double vn = a->rmsprop_rho * v[i] + (1.0 - a->rmsprop_rho) * phi * phi;
                v[i] = vn;
                double diff = 0.0;
                if (a->temperature > 1e-6) {
                    double xi;
                    if (nz->z_pred == NULL) xi = noise_normal(nz, NULL, i, tick, PURPOSE_SVLD0 + (uint32_t)(step - 1));
                    else xi = (nz->xi && step - 1 < nz->xi_rows) ? (double)nz->xi[(size_t)(step - 1) * n + i] : 0.0;
                    diff = sqrt(2.0 * eff * a->temperature) * xi;
                }
                hn[i] = clamp_logvol(h[i] + eff * phi / sqrt(vn + a->rmsprop_eps) + diff);
                if (last) kp[i] = kq;
            }
            memcpy(h, hn, (size_t)n * sizeof(double));
        }
    }

    /* phase 5: rejuvenation (:640-652) */
    if (a->use_rejuvenation && a->ksd_prev > a->rejuv_ksd_threshold && a->timestep > 10) {
        double gsd = sqrt(gvar > 1e-6 ? gvar : 1e-6);
        for (int i = 0; i < n; ++i) {
            double u = v2_uniform(nz, nz->u_rej, i, tick, PURPOSE_REJUV, 1.0);
            if (u < a->rejuv_prob) {
                double z = v2_normal(nz, nz->z_rej, i, tick, PURPOSE_REJUV, 0.0);
                h[i] = clamp_logvol((1.0 - a->rejuv_blend) * h[i] + a->rejuv_blend * (gmean + gsd * z));
            }
        }
    }

    /* phase 6: outputs (:656-707) */
    double ks = 0.0, sh = 0.0, sv = 0.0, mx = -1e10, sw = 0.0;
    for (int i = 0; i < n; ++i) {
        ks += kp[i];
        sh += h[i];
        sv += exp(h[i] * 0.5);
        if (log_w[i] > mx) mx = log_w[i];
    }
    for (int i = 0; i < n; ++i) sw += exp(log_w[i] - mx);
    double k2 = ks / ((double)n * (double)n);
    io->ksd = sqrt(k2 > 0.0 ? k2 : 0.0);
    io->h_mean = sh / n;
    io->vol = sv / n;
    io->loglik = mx + log((sw > 1e-10 ? sw : 1e-10) / n);
    if (a->use_adaptive_mu && a->timestep > 10)
        io->mu_state = clamp_d(a->mu_ema_alpha * io->h_mean + (1.0 - a->mu_ema_alpha) * io->mu_state, -10.0, 0.0);
    free(m);
    free(s);
    free(hs);
    free(hn);
    free(kp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* svpf_initialize's particle draw                                          */
/* ------------------------------------------------------------------------ */
/* src/svpf_optimized_graph.cu:73-97 (one normal per particle, clamped) and
 * :349-451 (stationary variance sigma^2/(1-rho^2+1e-6), times nu_s/(nu_s-2)
 * for the Student-t state with nu_s clamped to >= 2.5). */
void svpfo_init_particles(int n, double mu, double rho, double sigma_z, int student_t_state, double nu_state,
                          uint64_t seed, double* h) {
    double var = sigma_z * sigma_z / (1.0 - rho * rho + 1e-6);
    if (student_t_state) {
        double nus = nu_state > 2.5 ? nu_state : 2.5;
        var *= nus / (nus - 2.0);
    }
    double sd = sqrt(var);
    for (int i = 0; i < n; ++i) {
        uint32_t o[4];
        philox_draw(seed, 0u, 0u, PURPOSE_INIT, (uint32_t)i, o);
        h[i] = clamp_logvol(mu + sd * svpfo_normal_from_bits(o[0], o[1]));
    }
}

/* ------------------------------------------------------------------------ */
/* Smoothing of the reported h_mean (src/svpf_optimized_graph.cu:619-717)    */
/* ------------------------------------------------------------------------ */
double svpfo_smooth_step(svpfo_smooth* s, int timestep, double h_mean, double h_var, double y_t, double mu,
                         double rho, double sigma_z) {
    int k = s->lag;
    if (k > 8) k = 8;
    if (k < 1) k = 1;
    /* :643-653: store, advance the circular head */
    int head = s->head;
    s->h_mean[head] = h_mean;
    s->h_var[head] = h_var;
    s->y[head] = y_t;
    s->head = (head + 1) % k;
    if (timestep < k) return h_mean;
    /* :660-690: newest to oldest, each entry corrected from its (corrected) successor */
    double q = sigma_z * sigma_z;
    for (int lag = 1; lag < k; ++lag) {
        int ic = (s->head - lag - 1 + k) % k;
        int in = (s->head - lag + k) % k;
        double hc = s->h_mean[ic], hn = s->h_mean[in], vc = s->h_var[ic];
        double hp = mu + rho * (hc - mu);
        double pv = rho * rho * vc + q;
        double J = rho * vc / (pv + 1e-8);
        s->h_mean[ic] = hc + J * (hn - hp);
        s->h_var[ic] = vc * (1.0 - J * rho);
    }
    /* :701-717: the entry output_lag ticks back */
    int ol = s->output_lag;
    if (ol <= 0) return h_mean;
    if (ol >= k) ol = k - 1;
    return s->h_mean[(s->head - ol - 1 + k) % k];
}

/* ------------------------------------------------------------------------ */
/* Two-factor coordinate-wise step                                          */
/* ------------------------------------------------------------------------ */
/* include/svpf_two_factor.cuh:245-300: the factors' means and variances under
 * AR(1) dynamics; one scalar observation log y^2 = mu + f + s - offset + noise
 * (variance 4.93 + 2) whose innovation is split by the predicted variances. */
void svpfo_ekf_2f_update(const svpfo_2f_cfg* c, svpfo_2f_state* s, double y_t) {
    if (!s->guide_2f_initialized) {
        s->guide_mean_fast = 0.0;
        s->guide_mean_slow = 0.0;
        s->guide_var_fast = c->sigma_fast * c->sigma_fast / (1.0 - c->rho_fast * c->rho_fast);
        s->guide_var_slow = c->sigma_slow * c->sigma_slow / (1.0 - c->rho_slow * c->rho_slow);
        s->guide_2f_initialized = 1;
    }
    double mf = c->rho_fast * s->guide_mean_fast, ms = c->rho_slow * s->guide_mean_slow;
    double Pf = c->rho_fast * c->rho_fast * s->guide_var_fast + c->sigma_fast * c->sigma_fast;
    double Ps = c->rho_slow * c->rho_slow * s->guide_var_slow + c->sigma_slow * c->sigma_slow;
    double innov = log(y_t * y_t + 1e-8) + svpfo_implied_offset(c->nu) - (c->mu + mf + ms);
    double S = Pf + Ps + (4.93 + 2.0);
    double Kf = Pf / (S + 1e-8), Ks = Ps / (S + 1e-8);
    s->guide_mean_fast = mf + Kf * innov;
    s->guide_mean_slow = ms + Ks * innov;
    double vf = (1.0 - Kf) * Pf, vs = (1.0 - Ks) * Ps;
    s->guide_var_fast = vf > 1e-6 ? vf : 1e-6;
    s->guide_var_slow = vs > 1e-6 ? vs : 1e-6;
}

/* src/svpf_two_factor.cu:29-48 */
void svpfo_init_2f(int n, double rho_fast, double sigma_fast, double rho_slow, double sigma_slow, uint64_t seed,
                   double* hf, double* hs) {
    double sf = sigma_fast / sqrt(1.0 - rho_fast * rho_fast + 1e-8);
    double ss = sigma_slow / sqrt(1.0 - rho_slow * rho_slow + 1e-8);
    for (int i = 0; i < n; ++i) {
        uint32_t o[4];
        philox_draw(seed, 0u, 0u, PURPOSE_INIT_2F, (uint32_t)i, o);
        hf[i] = clamp_logvol(sf * svpfo_normal_from_bits(o[0], o[1]));
        hs[i] = clamp_logvol(ss * svpfo_normal_from_bits(o[2], o[3]));
    }
}

/* Student-t log density of y at log-variance h, as the two-factor gradient
 * kernel writes it (src/svpf_two_factor.cu:188-196): no clamp on exp(h). */
static double log_w_2f(double h, double y, double nu, double stc) {
    double ratio = y * y / (nu * exp(h) + 1e-8);
    return stc - 0.5 * h - 0.5 * (nu + 1.0) * log(1.0 + ratio);
}

/* src/svpf_opt_kernels.cu:527-599 on one coordinate: plain IMQ operator with the
 * reference sign, RMSProp, step eff, Langevin noise (Jacobi: all from x). */
static void transport_1d(int n, const double* x, const double* g, double* v, double bw, double eff, double temp,
                         double rms_rho, double rms_eps, int sign, uint64_t seed, uint32_t filter, uint32_t tick,
                         uint32_t purpose, double* xn) {
    double ib = 1.0 / (bw * bw), sgn = sign == 1 ? 1.0 : -1.0;
    for (int i = 0; i < n; ++i) {
        double ks = 0.0, gk = 0.0;
        for (int j = 0; j < n; ++j) {
            double d = x[i] - x[j], K = 1.0 / (1.0 + d * d * ib);
            ks += K * g[j];
            gk += sgn * 2.0 * d * ib * K * K;
        }
        double phi = (ks + gk) / n;
        double vn = rms_rho * v[i] + (1.0 - rms_rho) * phi * phi;
        v[i] = vn;
        double diff = 0.0;
        if (temp > 1e-6) {
            uint32_t o[4];
            philox_draw(seed, filter, tick, purpose, (uint32_t)i, o);
            diff = sqrt(2.0 * eff * temp) * svpfo_normal_from_bits(o[0], o[1]);
        }
        xn[i] = clamp_logvol(x[i] + eff * phi / sqrt(vn + rms_eps) + diff);
    }
}

/* Silverman bandwidth with a floor (src/svpf_two_factor.cu:256-313): one-pass variance. */
static double bw_2f(int n, const double* x, double floor_) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < n; ++i) { s1 += x[i]; s2 += x[i] * x[i]; }
    double m = s1 / n, var = s2 / n - m * m;
    double bw = 1.06 * sqrt(var > 1e-8 ? var : 1e-8) * pow((double)n, -0.2);
    return bw > floor_ ? bw : floor_;
}

int svpfo_tick_2f(const svpfo_2f_cfg* c, svpfo_2f_state* st, double* hf, double* hs, double* hfp, double* hsp,
                  double* vf, double* vs, double y_t, uint64_t seed, uint32_t filter_id, svpfo_2f_out* out) {
    const int n = c->n;
    const uint32_t tick = (uint32_t)st->timestep;
    if (n < 1 || c->n_stein_steps < 1) return 1;
    double* lw = (double*)calloc((size_t)n, sizeof(double));
    double* gf = (double*)calloc((size_t)n, sizeof(double));
    double* gs = (double*)calloc((size_t)n, sizeof(double));
    double* xn = (double*)calloc((size_t)n, sizeof(double));
    if (!lw || !gf || !gs || !xn) return -1;
    const double stc = svpfo_student_t_const(c->nu);

    /* predict (src/svpf_two_factor.cu:55-93): AR(1) of each mean-zero factor, MIM on the fast one */
    for (int i = 0; i < n; ++i) {
        uint32_t o[4];
        philox_draw(seed, filter_id, tick, PURPOSE_PREDICT, (uint32_t)i, o);
        double zf = svpfo_normal_from_bits(o[0], o[1]);
        double u = svpfo_u01_from_bits(o[2]);
        philox_draw(seed, filter_id, tick, PURPOSE_PREDICT_SLOW, (uint32_t)i, o);
        double zs = svpfo_normal_from_bits(o[0], o[1]);
        if (c->use_mim && u < c->mim_jump_prob) zf *= c->mim_jump_scale;
        hfp[i] = hf[i];
        hsp[i] = hs[i];
        hf[i] = clamp_logvol(c->rho_fast * hfp[i] + c->sigma_fast * zf);
        hs[i] = clamp_logvol(c->rho_slow * hsp[i] + c->sigma_slow * zs);
    }
    /* predictive log-likelihood of the predicted particles (before the guide) */
    for (int i = 0; i < n; ++i) lw[i] = log_w_2f(c->mu + hf[i] + hs[i], y_t, c->nu, stc);
    out->pred_loglik = lse_mean(lw, n);

    /* guide (src/svpf_optimized_graph.cu:780-807, src/svpf_two_factor.cu:320-360): the
     * factors' means are the previous tick's */
    if (c->use_guide) {
        svpfo_ekf_2f_update(c, st, y_t);
        double g = c->guide_strength, cf = st->mean_fast_prev, cs = st->mean_slow_prev;
        for (int i = 0; i < n; ++i) {
            if (c->use_guide_preserving) {
                hf[i] = clamp_logvol((1.0 - g) * cf + g * st->guide_mean_fast + (hf[i] - cf));
                hs[i] = clamp_logvol((1.0 - g) * cs + g * st->guide_mean_slow + (hs[i] - cs));
            } else {
                hf[i] = clamp_logvol((1.0 - g) * hf[i] + g * st->guide_mean_fast);
                hs[i] = clamp_logvol((1.0 - g) * hs[i] + g * st->guide_mean_slow);
            }
        }
    }

    /* bandwidths (src/svpf_two_factor.cu:256-313) */
    double bwf = bw_2f(n, hf, c->bw_floor_fast), bws = bw_2f(n, hs, c->bw_floor_slow);

    /* Stein loop (src/svpf_optimized_graph.cu:822-868): beta = (a+1)/A, step 0.5/S x beta,
     * gradient then the 1-D transport of each factor */
    const int na = c->use_annealing ? c->n_anneal_steps : 1;
    const double base_step = 0.5 / (double)c->n_stein_steps;
    const double temp = c->use_svld ? c->temperature : 0.0;
    const double s2f = c->sigma_fast * c->sigma_fast + 1e-8, s2s = c->sigma_slow * c->sigma_slow + 1e-8;
    const double wf = s2f / (s2f + s2s), ws = 1.0 - wf;
    int it = 0;
    for (int a = 0; a < na; ++a) {
        const double beta = (double)(a + 1) / (double)na;
        for (int k = 0; k < c->n_stein_steps; ++k, ++it) {
            /* src/svpf_two_factor.cu:171-249 */
            for (int i = 0; i < n; ++i) {
                double h = c->mu + hf[i] + hs[i];
                double ratio = y_t * y_t / (c->nu * exp(h) + 1e-8);
                lw[i] = log_w_2f(h, y_t, c->nu, stc);
                double gl = c->use_exact_gradient ? -0.5 + 0.5 * (c->nu + 1.0) * ratio / (1.0 + ratio + 1e-8)
                                                  : 0.5 * (log(y_t * y_t + 1e-8) - h);
                gf[i] = -(hf[i] - c->rho_fast * hfp[i]) / s2f + beta * wf * gl;
                gs[i] = -(hs[i] - c->rho_slow * hsp[i]) / s2s + beta * ws * gl;
            }
            const double eff = base_step * beta;
            transport_1d(n, hf, gf, vf, bwf, eff, temp, c->rmsprop_rho, c->rmsprop_eps, c->stein_sign, seed,
                         filter_id, tick, PURPOSE_SVLD0 + (uint32_t)it, xn);
            memcpy(hf, xn, (size_t)n * sizeof(double));
            transport_1d(n, hs, gs, vs, bws, eff, temp, c->rmsprop_rho, c->rmsprop_eps, c->stein_sign, seed,
                         filter_id, tick, PURPOSE_SVLD_SLOW0 + (uint32_t)it, xn);
            memcpy(hs, xn, (size_t)n * sizeof(double));
        }
    }

    /* outputs (src/svpf_two_factor.cu:394-468) */
    double sf = 0.0, ss = 0.0, sv = 0.0;
    for (int i = 0; i < n; ++i) {
        sf += hf[i];
        ss += hs[i];
        sv += exp(0.5 * (c->mu + hf[i] + hs[i]));
    }
    out->loglik = lse_mean(lw, n);
    out->mean_fast = sf / n;
    out->mean_slow = ss / n;
    out->h_mean = c->mu + out->mean_fast + out->mean_slow;
    out->vol = sv / n;
    out->bw_fast = bwf;
    out->bw_slow = bws;
    out->steps = it;
    st->mean_fast_prev = out->mean_fast;
    st->mean_slow_prev = out->mean_slow;
    st->vol_prev = out->vol;
    st->h_mean_prev = out->h_mean;
    st->timestep++;
    free(lw); free(gf); free(gs); free(xn);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Joint state-parameter filter (svpf_joint/)                               */
/* ------------------------------------------------------------------------ */
static double sigm(double x) { return 1.0 / (1.0 + exp(-x)); }
/* svpf_joint_kernels.cu:28-30: exp with its argument clamped to [-20, 20] */
static double jsafe_exp(double x) { return exp(clamp_d(x, -20.0, 20.0)); }

/* svpf_joint.cu:138-171, svpf_joint_kernels.cu:47-73 */
void svpfo_joint_init(const svpfo_joint_cfg* c, uint64_t seed, double* x, double* hp) {
    const int n = c->n;
    for (int i = 0; i < n; ++i) {
        uint32_t a[4], b[4];
        philox_draw(seed, 0u, 0u, PURPOSE_JOINT_INIT_A, (uint32_t)i, a);
        philox_draw(seed, 0u, 0u, PURPOSE_JOINT_INIT_B, (uint32_t)i, b);
        x[i] = c->mu_prior_mean + 0.5 * svpfo_normal_from_bits(a[0], a[1]);
        x[n + i] = c->mu_prior_mean + 0.2 * svpfo_normal_from_bits(a[2], a[3]);
        x[2 * n + i] = c->rho_prior_mean + 0.2 * svpfo_normal_from_bits(b[0], b[1]);
        x[3 * n + i] = c->sigma_prior_mean + 0.2 * svpfo_normal_from_bits(b[2], b[3]);
        hp[i] = x[i];
    }
}

/* svpf_joint_kernels.cu:134-320 */
void svpfo_joint_grad(const svpfo_joint_cfg* c, const double xj[4], double h_prev, double y_t, double g[4],
                      double* log_w) {
    const double h = xj[0], mt = xj[1], rt = xj[2], st = xj[3];
    const double mu = mt, rho = sigm(rt);
    double sigma = jsafe_exp(st);
    if (sigma < 0.01) sigma = 0.01;
    const double var = sigma * sigma;
    const double ysq = y_t * y_t;
    const double A = ysq / (jsafe_exp(h) + 1e-8) / c->nu;
    *log_w = c->student_t_const - 0.5 * h - (c->nu + 1.0) * 0.5 * log1p(A > -0.999 ? A : -0.999);
    const double gh_lik = (log(ysq + 1e-10) - h + c->lik_offset) / 1.4;
    /* observation surprise against the previous h */
    const double z2o = ysq / (jsafe_exp(h_prev) + 1e-8);
    const double surprise = z2o - 4.0 > 0.0 ? z2o - 4.0 : 0.0;
    const double inflate = 1.0 + 0.3 * (surprise < 20.0 ? surprise : 20.0);
    double vm = var * inflate;
    if (vm < 0.03 * 0.03) vm = 0.03 * 0.03;
    const int panic = z2o > 9.0;
    if (panic && vm < 1.0) vm = 1.0;
    const double tw = panic ? 0.3 : 1.0;
    const double diff = h - (mu + rho * (h_prev - mu));
    const double gh_tr = tw * (-diff / vm);
    const double gm_tr = tw * (diff / vm) * (1.0 - rho);
    const double gr_tr = tw * (diff / vm) * (h_prev - mu) * rho * (1.0 - rho);
    const double gs_tr = diff * diff / var - 1.0;
    double fobs = 0.0;
    if (z2o > 9.0) fobs = (z2o - 9.0) * 0.5 < 5.0 ? (z2o - 9.0) * 0.5 : 5.0;
    else if (z2o < 1.5) fobs = -(1.5 - z2o);
    const double mr = -0.2 * (st - c->sigma_prior_mean);
    const double pm = -(mt - c->mu_prior_mean) / c->mu_prior_var;
    const double pr = -(rt - c->rho_prior_mean) / c->rho_prior_var;
    const double ps = -(st - c->sigma_prior_mean) / c->sigma_prior_var;
    g[0] = clamp_d(gh_lik + gh_tr, -10.0, 10.0);
    g[1] = clamp_d(gm_tr + c->prior_weight * pm, -5.0, 5.0);
    g[2] = clamp_d(gr_tr + c->prior_weight * pr, -5.0, 5.0);
    g[3] = clamp_d(gs_tr + fobs + mr + c->prior_weight * ps, -5.0, panic ? 10.0 : 5.0);
}

int svpfo_joint_step(const svpfo_joint_cfg* c, double* x, double* hp, double y_t, uint64_t seed, uint32_t tick,
                     svpfo_joint_out* out) {
    const int n = c->n;
    double* g = (double*)malloc((size_t)4 * n * sizeof(double));
    double* xn = (double*)malloc((size_t)4 * n * sizeof(double));
    double* lw = (double*)malloc((size_t)n * sizeof(double));
    if (!g || !xn || !lw) return -1;
    /* predict with parameter diffusion (svpf_joint_kernels.cu:79-128) */
    for (int i = 0; i < n; ++i) {
        uint32_t a[4], b[4];
        philox_draw(seed, 0u, tick, PURPOSE_JOINT_A, (uint32_t)i, a);
        philox_draw(seed, 0u, tick, PURPOSE_JOINT_B, (uint32_t)i, b);
        const double h = x[i];
        hp[i] = h;
        const double mt = clamp_d(x[n + i] + c->diffusion_mu * svpfo_normal_from_bits(a[0], a[1]), -10.0, 5.0);
        const double rt = clamp_d(x[2 * n + i] + c->diffusion_rho * svpfo_normal_from_bits(a[2], a[3]), -5.0, 5.0);
        const double st = clamp_d(x[3 * n + i] + c->diffusion_sigma * svpfo_normal_from_bits(b[0], b[1]), -5.0, 2.0);
        x[n + i] = mt;
        x[2 * n + i] = rt;
        x[3 * n + i] = st;
        double sigma = jsafe_exp(st);
        if (sigma < 0.01) sigma = 0.01;
        const double rho = sigm(rt);
        x[i] = clamp_logvol(mt + rho * (h - mt) + sigma * svpfo_normal_from_bits(b[2], b[3]));
    }
    /* bandwidths (svpf_joint_kernels.cu:620-663): sqrt(2 var / ln(n+1)), clamped */
    double ib[4];
    for (int d = 0; d < 4; ++d) {
        double s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < n; ++i) { s1 += x[d * n + i]; s2 += x[d * n + i] * x[d * n + i]; }
        const double m = s1 / n, var = s2 / n - m * m;
        double bsq = 2.0 * var / log((double)n + 1.0);
        if (bsq < 1e-6) bsq = 1e-6;
        const double bw = clamp_d(sqrt(bsq), 0.01, 2.0);
        out->bw[d] = bw;
        ib[d] = 1.0 / (bw * bw + 1e-8);
    }
    /* With no Stein iteration the log weights are those of the predicted particles
     * (the reference's diagnostics read a buffer no kernel has written then). */
    if (c->n_stein_steps < 1)
        for (int j = 0; j < n; ++j) {
            double xj[4] = {x[j], x[n + j], x[2 * n + j], x[3 * n + j]};
            svpfo_joint_grad(c, xj, hp[j], y_t, &g[4 * j], &lw[j]);
        }
    /* Stein iterations (svpf_joint.cu:207-276, svpf_joint_kernels.cu:326-470).  Every
     * particle reads the positions of the previous iteration (Jacobi): the reference's
     * kernel stages all positions in shared memory before its update, which is this
     * reading when its blocks do not interleave. */
    for (int s = 0; s < c->n_stein_steps; ++s) {
        for (int j = 0; j < n; ++j) {
            double xj[4] = {x[j], x[n + j], x[2 * n + j], x[3 * n + j]};
            svpfo_joint_grad(c, xj, hp[j], y_t, &g[4 * j], &lw[j]);
        }
        for (int j = 0; j < n; ++j) {
            const double zo = y_t * y_t / (exp(hp[j]) + 1e-8);
            const int sur = zo > 9.0;
            const double eff[4] = {c->step_h * (sur ? 5.0 : 1.0), c->step_mu * (sur ? 2.0 : 1.0), c->step_rho,
                                   c->step_sigma * (sur ? 4.0 : 1.0)};
            double phi[4] = {0.0, 0.0, 0.0, 0.0}, sk = 0.0;
            for (int i = 0; i < n; ++i) {
                double d[4], dist = 0.0;
                for (int q = 0; q < 4; ++q) {
                    d[q] = x[q * n + i] - x[q * n + j];
                    dist += d[q] * d[q] * ib[q];
                }
                const double K = 1.0 / (1.0 + dist);
                sk += K;
                for (int q = 0; q < 4; ++q) phi[q] += K * g[4 * i + q] - 2.0 * d[q] * ib[q] * K * K;
            }
            double norm = sk < (double)n ? sk : (double)n;
            if (norm < 32.0) norm = 32.0;
            xn[j] = clamp_logvol(x[j] + eff[0] * phi[0] / norm);
            xn[n + j] = clamp_d(x[n + j] + eff[1] * phi[1] / norm, -10.0, 5.0);
            xn[2 * n + j] = clamp_d(x[2 * n + j] + eff[2] * phi[2] / norm, -5.0, 5.0);
            xn[3 * n + j] = clamp_d(x[3 * n + j] + eff[3] * phi[3] / norm, -5.0, 2.0);
        }
        memcpy(x, xn, (size_t)4 * n * sizeof(double));
    }
    /* extract (svpf_joint_kernels.cu:476-614) and the host diagnostics (svpf_joint.cu:280-345) */
    double sh = 0.0, sm = 0.0, sr = 0.0, ss = 0.0, qm = 0.0, qr = 0.0, qs = 0.0;
    double tm = 0.0, tr = 0.0, ts = 0.0, um = 0.0, ur = 0.0, us = 0.0;
    for (int i = 0; i < n; ++i) {
        const double mt = x[n + i], rt = x[2 * n + i], st = x[3 * n + i];
        const double r = sigm(rt), sg = exp(st);
        sh += x[i];
        sm += mt; sr += r; ss += sg;
        qm += mt * mt; qr += r * r; qs += sg * sg;
        tm += mt; tr += rt; ts += st;
        um += mt * mt; ur += rt * rt; us += st * st;
    }
    const double inv = 1.0 / n;
    out->h_mean = sh * inv;
    out->vol_mean = exp(0.5 * out->h_mean);
    out->mu_mean = sm * inv;
    out->rho_mean = sr * inv;
    out->sigma_mean = ss * inv;
#define SDV(q, m) sqrt(((q) * inv - (m) * (m)) > 0.0 ? ((q) * inv - (m) * (m)) : 0.0)
    out->mu_std = SDV(qm, out->mu_mean);
    out->rho_std = SDV(qr, out->rho_mean);
    out->sigma_std = SDV(qs, out->sigma_mean);
    out->std_mu_tilde = SDV(um, tm * inv);
    out->std_rho_tilde = SDV(ur, tr * inv);
    out->std_sigma_tilde = SDV(us, ts * inv);
#undef SDV
    out->mu_collapsed = out->std_mu_tilde < c->collapse_thresh_mu;
    out->rho_collapsed = out->std_rho_tilde < c->collapse_thresh_rho;
    out->sigma_collapsed = out->std_sigma_tilde < c->collapse_thresh_sigma;
    double mx = lw[0];
    for (int i = 1; i < n; ++i) if (lw[i] > mx) mx = lw[i];
    double sw = 0.0, sw2 = 0.0;
    for (int i = 0; i < n; ++i) { double w = exp(lw[i] - mx); sw += w; sw2 += w * w; }
    out->ess = sw * sw / (sw2 + 1e-10);
    out->log_likelihood = mx + log(sw) - log((double)n);
    free(g); free(xn); free(lw);
    return 0;
}
