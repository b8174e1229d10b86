// svpf_types.h -- plain-old-data blocks shared by the HIP kernels and the host
// layer.  No HIP or torch types: the same header is read by the C-ABI host code
// and by the kernels.
//
// The per-filter parameter block carries the SVPFState feature flags and model
// parameters that the reference's step reads (include/svpf.cuh:290-547,
// src/svpf_optimized_graph.cu:915-1389), flattened so one block describes one
// filter of a batch.
#ifndef SVPF_TYPES_H
#define SVPF_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Feature flags (svpf_fparams.flags).
enum {
    SVPF_F_EXACT_GRAD = 1u << 0,
    SVPF_F_NEWTON = 1u << 1,
    SVPF_F_FULL_NEWTON = 1u << 2,
    SVPF_F_SVLD = 1u << 3,
    SVPF_F_ANNEAL = 1u << 4,
    SVPF_F_ADAPT_BETA = 1u << 5,
    SVPF_F_LINEAR_BETA = 1u << 6,
    SVPF_F_MIM = 1u << 7,
    SVPF_F_ASYM_RHO = 1u << 8,
    SVPF_F_LOCAL_PARAMS = 1u << 9,
    SVPF_F_GUIDED = 1u << 10,
    SVPF_F_GUIDE = 1u << 11,
    SVPF_F_GUIDE_PRESERVING = 1u << 12,
    SVPF_F_ADAPT_GUIDE = 1u << 13,
    SVPF_F_REJUV = 1u << 14,
    SVPF_F_FAN = 1u << 15,
    SVPF_F_ADAPT_MU = 1u << 16,
    SVPF_F_ADAPT_SIGMA = 1u << 17,
    SVPF_F_KSD_EARLY_STOP = 1u << 18,
    SVPF_F_PAPER_SIGN = 1u << 19,
    SVPF_F_HEUN = 1u << 20,           // Heun predictor-corrector Stein step
    SVPF_F_STUDENT_STATE = 1u << 21,  // Student-t state noise and transition prior (nu_state)
    SVPF_F_TWO_FACTOR = 1u << 22,     // the coordinate-wise two-factor step (its own kernel)
};

// Bandwidth rules.
enum { SVPF_BW_REFERENCE_EMA = 0, SVPF_BW_MEDIAN = 1, SVPF_BW_SILVERMAN = 2 };

// Per-filter constant parameters (64 x 4 bytes).
typedef struct svpf_fparams {
    int32_t n;
    uint32_t flags;
    int32_t n_anneal;
    int32_t stein_min;
    int32_t stein_max;
    int32_t bandwidth_mode;
    int32_t pad0[2];
    float mu, rho, sigma_z, gamma;
    float nu, lik_offset, temperature, rms_rho;
    float rms_eps, mim_p, mim_scale, rho_up;
    float rho_down, delta_rho, delta_sigma, ga_base;
    float ga_shock, ga_thr, gs_base, gs_max;
    float gs_thr, rej_thr, rej_p, rej_blend;
    float mu_q, mu_rscale, mu_min, mu_max;
    float sb_thr, sb_max, ksd_thr, st_const;
    float implied_offset, base_step;
    float nu_state;
    int32_t coop_spin_limit;  // polls before a cooperative-kernel exchange times out (0: 2^20;
                              // negative: a forced time-out after that many polls)
    // the coordinate-wise two-factor step (SVPF_F_TWO_FACTOR)
    float rho_fast, sigma_fast, rho_slow, sigma_slow;
    float bw_floor_fast, bw_floor_slow, guide_strength;
    int32_t n_stein;          // two-factor step: Stein iterations per annealing stage
    int32_t coop_fault_wg;    // failure-path testing: this cooperative workgroup's final-exchange waits
                              // give up (-1: none)
    float pad1[11];
} svpf_fparams;

// Per-filter scalar state carried between ticks (host fields of SVPFState and
// the device scalars of SVPFOptimizedState).  32 x 4 bytes.
typedef struct svpf_fstate {
    int32_t timestep;
    int32_t guide_initialized;
    int32_t stein_steps_used;
    int32_t status;            // bit 0: the kernel rejected this filter; bit 1: a tick's wait for
                               // the filter's previous tick hit the spin limit
    float ksd_prev, vol_prev, h_mean_prev, guide_mean;
    float guide_var, bw_sq_ema, return_ema, return_var;
    float mu_state, mu_var, sigma_z_eff, y_prev;
    float last_loglik, last_vol, last_h_mean, last_bw;
    float last_ksd, last_pred_loglik, last_guide_strength;
    uint32_t epoch;            // last exchange epoch of the cooperative kernel
    float med_r2;              // last median pairwise r^2 (seeds the next tick's selection)
    // two-factor step: the factors' EKF guide and their particle means after the last tick
    float g2f_mean_fast, g2f_mean_slow, g2f_var_fast, g2f_var_slow;
    float mean_fast_prev, mean_slow_prev;
    int32_t g2f_initialized;
} svpf_fstate;

// Per-tick output record (8 floats): written to out[t * SVPF_OUTW + k].
enum {
    SVPF_O_LOGLIK = 0, SVPF_O_VOL = 1, SVPF_O_HMEAN = 2, SVPF_O_KSD = 3,
    SVPF_O_BW = 4, SVPF_O_PRED_LL = 5, SVPF_O_STEPS = 6, SVPF_O_BWSQ = 7, SVPF_OUTW = 8
};

// Per-filter descriptor for one launch.
typedef struct svpf_fdesc {
    int64_t off;               // first particle of this filter in the particle pool
    const float* y;            // observations, y[t] for t in [0, T)
    float* out;                // per-tick records, out[t * SVPF_OUTW + k] (may be null)
    const float* z_pred;       // pre-drawn predict normals [T][n], null: counter generator
    const float* u_pred;       // pre-drawn MIM selector uniforms [T][n] (may be null)
    const float* xi;           // pre-drawn Langevin normals [T][xi_rows][n]
    const float* u_rej;        // pre-drawn rejuvenation uniforms [T][n] (may be null)
    const float* z_rej;        // pre-drawn rejuvenation normals [T][n] (may be null)
    float* log_w;              // final log weights [n] written at the end of a launch (may be null)
    float* h_pred;             // particles after predict + guide of each tick [n] (may be null)
    uint64_t seed;             // counter-generator key
    int32_t xi_rows;
    int32_t param_idx;         // index into the svpf_fparams array
    uint32_t filter_id;        // counter-generator filter word
    int32_t tile;              // tile-shape index chosen for this filter's n
} svpf_fdesc;

// Noise purposes: third word of the counter-generator counter.
// SVPF_PURPOSE_STATE_T0 + k: normal k of the chi-squared sum of a Student-t state draw.
// Two-factor step: the fast factor's predict normal and MIM uniform come from
// SVPF_PURPOSE_PREDICT, the slow factor's normal from SVPF_PURPOSE_PREDICT_SLOW;
// the Langevin normals of iteration k from SVPF_PURPOSE_SVLD0 + k (fast) and
// SVPF_PURPOSE_SVLD_SLOW0 + k (slow); its initial draw from SVPF_PURPOSE_INIT_2F.
// Joint state-parameter filter: the diffusion normals of mu~ and rho~ from
// SVPF_PURPOSE_JOINT_A, of sigma~ and h from SVPF_PURPOSE_JOINT_B; its initial
// draw of (h, mu~) from SVPF_PURPOSE_JOINT_INIT_A and of (rho~, sigma~) from
// SVPF_PURPOSE_JOINT_INIT_B.
enum { SVPF_PURPOSE_PREDICT = 0, SVPF_PURPOSE_REJUV = 1, SVPF_PURPOSE_PREDICT_SLOW = 2, SVPF_PURPOSE_INIT = 255,
       SVPF_PURPOSE_INIT_2F = 254, SVPF_PURPOSE_SVLD0 = 16, SVPF_PURPOSE_SVLD_SLOW0 = 1024,
       SVPF_PURPOSE_STATE_T0 = 128, SVPF_PURPOSE_JOINT_A = 3, SVPF_PURPOSE_JOINT_B = 4,
       SVPF_PURPOSE_JOINT_INIT_A = 253, SVPF_PURPOSE_JOINT_INIT_B = 252 };

// Joint state-parameter filter: the step kernel's parameters (passed by value).
typedef struct svpf_jparams {
    int32_t n, n_stein;
    float step_h, step_mu, step_rho, step_sigma;
    float diff_mu, diff_rho, diff_sigma;
    float thr_mu, thr_rho, thr_sigma;
    float mu_pm, mu_pv, rho_pm, rho_pv, sig_pm, sig_pv;
    float nu, st_const, lik_offset, prior_weight;
    uint32_t k0, k1;
    int32_t filter, pad;
} svpf_jparams;
// Words of one joint-filter step record: the reference's diagnostics structure
// (11 floats, 3 ints, 2 floats) followed by the four bandwidths.
#define SVPF_JOINT_RECW_DEV 20

// ---------------------------------------------------------------------------
// Differentiable form: B filters of n particles with the plain feature set
// (predict, median IMQ bandwidth, S Stein iterations with beta 0.3 -> 1.0,
// mixture prior, exact Student-t score, local Newton, RMSProp, SVLD), per
// filter theta = (mu, rho, sigma_z, nu).
// ---------------------------------------------------------------------------

// Configuration shared by every filter of a launch (24 x 4 bytes).
typedef struct svpf_dconfig {
    int32_t n;            // particles per filter
    int32_t S;            // Stein iterations per tick
    int32_t T;            // ticks
    int32_t C;            // checkpoint interval (ticks)
    int32_t paper_sign;   // 1: repulsive kernel gradient, 0: the reference default
    int32_t pad0[3];
    float gamma, lik_offset, temperature, rms_rho;
    float rms_eps, pad1[3];
    float pad2[8];
} svpf_dconfig;

// Per-filter descriptor.  All pointers are device pointers; optional ones may be null.
typedef struct svpf_ddesc {
    const float* theta;   // [4] mu, rho, sigma_z, nu
    const float* y;       // [T] observations (y_prev of tick 0 is 0)
    const float* h0;      // [n] initial particles (RMSProp state starts at 0)
    const float* z;       // [T][n] pre-drawn predict normals, null: counter generator
    const float* xi;      // [T][S][n] pre-drawn Langevin normals (with z)
    float* pred_ll;       // [T] predictive log-likelihood per tick (forward output)
    float* ell2;          // [T] median pairwise r^2 per tick (forward output, read by the backward)
    float* ckpt;          // [ceil(T/C)][2][n] checkpoints (x, v) at ticks 0, C, 2C, ...
    float* tape;          // [C] tick slots of svpf_diff_slot_floats(n, S) floats (backward scratch)
    const float* gpl;     // [T] upstream gradient dL/d pred_ll (backward input)
    float* gtheta;        // [4] dL/d theta (backward output)
    float* gh0;           // [n] dL/d h0 (backward output, optional)
    float* hT;            // [n] particles after the last tick (forward output, optional)
    uint64_t seed;        // counter-generator key
    uint32_t filter_id;
    int32_t pad;
    float* drift;         // [T][S][n] Stein drift phi of every iteration (forward output, optional): the
                          // backward's recompute replays the positions from it instead of redoing the transport
    int32_t* mpair;       // [T][2] the median pair (a, b) the adjoint differentiates through, -1 where the
                          // bandwidth is clamped (written where the tape is recorded; optional)
} svpf_ddesc;

// ---------------------------------------------------------------------------
// Two-factor (2-D) differentiable form: particle x = (f, s), log-vol
// h = mu + f + s, per-filter theta = (mu, rho_f, sigma_f, rho_s, sigma_s, nu).
// Arrays of particles are [2][n] (f slab, then s slab).
// ---------------------------------------------------------------------------

// Configuration shared by every filter of a launch (20 x 4 bytes).
typedef struct svpf_d2config {
    int32_t n;            // particles per filter
    int32_t S;            // Stein iterations per tick
    int32_t T;            // ticks
    int32_t C;            // checkpoint interval (ticks)
    int32_t paper_sign;   // 1: repulsive kernel gradient, 0: the reference default
    int32_t pad0[3];
    float lik_offset, temperature, rms_rho, rms_eps;
    float pad1[8];
} svpf_d2config;

// Per-tick output record of the two-factor forward (8 floats).
enum { SVPF_O2_PRED_LL = 0, SVPF_O2_HMEAN = 1, SVPF_O2_FMEAN = 2, SVPF_O2_SMEAN = 3, SVPF_O2_VOL = 4,
       SVPF_O2_BWF = 5, SVPF_O2_BWS = 6, SVPF_O2W = 8 };

// Per-filter descriptor (device pointers; optional ones may be null).
typedef struct svpf_d2desc {
    const float* theta;   // [6] mu, rho_f, sigma_f, rho_s, sigma_s, nu
    const float* y;       // [T]
    const float* x0;      // [2][n] initial factors (RMSProp state starts at 0)
    const float* z;       // [T][2][n] pre-drawn predict normals, null: counter generator
    const float* xi;      // [T][S][2][n] pre-drawn Langevin normals (with z)
    float* out;           // [T][SVPF_O2W] per-tick records (forward output)
    float* ell2;          // [T][2] median pairwise r^2 per coordinate (forward output, read by the backward)
    float* ckpt;          // [ceil(T/C)][4][n] checkpoints (f, s, v_f, v_s)
    float* tape;          // [C] tick slots of svpf_diff2d_slot_floats(n, S) floats
    const float* gpl;     // [T] dL/d pred_ll (backward input)
    float* gtheta;        // [6] dL/d theta (backward output)
    float* gx0;           // [2][n] dL/d x0 (backward output, optional)
    float* xT;            // [2][n] factors after the last tick (forward output, optional)
    uint64_t seed;
    uint32_t filter_id;
    int32_t pad;
    float* drift;         // [T][S][3][n] per iteration the Stein drift (f, s) and the mixture prior's shift
                          // (forward output, optional): the backward's recompute replays the positions from it
                          // instead of redoing the transport, and skips the prior's nearest-component pass
    int32_t* mpair;       // [T][2][2] per coordinate the median pair (a, b) the adjoint differentiates
                          // through, -1 where that bandwidth is clamped (optional)
} svpf_d2desc;

#ifdef __cplusplus
}
#endif
#endif
