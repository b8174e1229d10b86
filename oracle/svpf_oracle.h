/*
 * svpf_oracle.h -- scalar FP64 restatement of the SVPF per-tick step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product path links
 * or calls it.
 *
 * Every function restates the reference algorithm in double precision and
 * cites the reference file:line it follows (paths relative to the reference
 * repository root).
 */
#ifndef SVPF_ORACLE_H
#define SVPF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bandwidth rules (the reference has three; see svpfo_bandwidth). */
enum { SVPFO_BW_REFERENCE_EMA = 0, SVPFO_BW_MEDIAN = 1, SVPFO_BW_SILVERMAN = 2 };

/* Per-filter configuration: the SVPFState feature flags and parameters that the
 * step reads (include/svpf.cuh:290-547) plus the three extensions the
 * benchmark configurations need (bandwidth_mode, use_linear_beta,
 * use_ksd_early_stop). */
typedef struct svpfo_cfg {
    int n;
    int bandwidth_mode;
    int use_exact_gradient;
    int use_newton;
    int use_full_newton;
    int stein_sign;            /* 0 legacy (attract), 1 paper (repel) */
    int use_svld;
    int use_annealing;
    int n_anneal_steps;
    int use_adaptive_beta;
    int use_linear_beta;       /* beta_k = 0.3 + 0.7 k/(K-1) */
    int stein_min_steps;
    int stein_max_steps;
    int use_mim;
    int use_asymmetric_rho;
    int use_local_params;
    int use_guided;
    int use_guide;
    int use_guide_preserving;
    int use_adaptive_guide;
    int use_rejuvenation;
    int use_fan_mode;
    int use_adaptive_mu;
    int use_adaptive_sigma;
    int use_ksd_early_stop;
    int use_heun;              /* Heun predictor-corrector Stein step */
    int use_student_t_state;   /* Student-t state noise and transition prior */
    int pad_;
    double mu, rho, sigma_z, gamma;
    double nu, lik_offset;
    double temperature, rmsprop_rho, rmsprop_eps;
    double mim_jump_prob, mim_jump_scale;
    double rho_up, rho_down, delta_rho, delta_sigma;
    double guided_alpha_base, guided_alpha_shock, guided_innovation_threshold;
    double guide_strength_base, guide_strength_max, guide_innovation_threshold;
    double rejuv_ksd_threshold, rejuv_prob, rejuv_blend;
    double mu_process_var, mu_obs_var_scale, mu_min, mu_max;
    double sigma_boost_threshold, sigma_boost_max;
    double ksd_improvement_threshold;
    double nu_state;           /* Student-t state degrees of freedom */
} svpfo_cfg;

/* Scalar filter state carried between ticks (host fields of SVPFState and the
 * device scalars of SVPFOptimizedState). */
typedef struct svpfo_state {
    int timestep;
    int guide_initialized;
    int stein_steps_used;
    int pad_;
    double ksd_prev;
    double vol_prev;
    double h_mean_prev;       /* d_h_mean_prev: mean of h after the last tick */
    double guide_mean, guide_var;
    double bw_sq_ema, return_ema, return_var;
    double mu_state, mu_var;
    double sigma_z_effective;
} svpfo_state;

/* Pre-drawn noise for one tick.  Any pointer may be NULL when the matching
 * feature is off.  xi holds xi_rows rows of n Langevin normals; row s feeds
 * Stein iteration s of the tick.  When z_pred is NULL the noise is drawn from
 * the counter-based generator (svpfo_philox_*) keyed by (seed, filter, tick). */
typedef struct svpfo_noise {
    const float* z_pred;
    const float* u_pred;
    const float* xi;
    int xi_rows;
    int pad_;
    const float* u_rej;
    const float* z_rej;
    uint64_t seed;
    uint32_t filter_id;
    uint32_t pad2_;
} svpfo_noise;

/* Per-tick outputs. */
typedef struct svpfo_out {
    double loglik;        /* reference loglik: logmeanexp of log_w at the last gradient evaluation */
    double pred_loglik;   /* predictive loglik: logmeanexp of log p(y_t | h_pred) before any transport */
    double vol;           /* mean exp(h/2) */
    double h_mean;        /* mean h */
    double ksd;           /* KSD of the last Stein iteration */
    double bandwidth;     /* bw used by the transport this tick */
    double bw_sq;         /* raw ell^2: median r^2 / EMA state / Silverman bw^2 (before the bw clamp) */
    int steps;            /* Stein iterations executed */
    int pad_;
} svpfo_out;

/* ---- closed-form helpers (each exported for the known-answer tests) ---- */
double svpfo_student_t_const(double nu);
double svpfo_implied_offset(double nu);
double svpfo_log_w(double h, double y, double nu, double st_const);
double svpfo_median_r2(const double* h, int n);
float  svpfo_median_r2_f32(const float* h, int n);
double svpfo_adaptive_beta(double ksd_prev, int anneal_idx, int n_anneal);
int    svpfo_stein_budget(const svpfo_cfg* c, const svpfo_state* s);
void   svpfo_ekf_update(const svpfo_cfg* c, svpfo_state* s, double y_t, double mu, double rho, double sigma_z);

/* Mixture-prior + likelihood gradient for all particles (O(N^2)). */
void svpfo_gradient(const svpfo_cfg* c, const double* h, const double* h_prev, double y_t, double y_prev,
                    double mu, double sigma_z, double beta,
                    double* grad, double* log_w, double* precond, double* curv);

/* Stein operator phi and the reference KSD partial sums for all particles. */
void svpfo_stein_phi(const svpfo_cfg* c, const double* h, const double* s, const double* curv,
                     double inv_bw_sq, double* phi, double* ksd_partial);

/* One full tick.  h, h_prev, v are length-n working arrays updated in place. */
int svpfo_tick(const svpfo_cfg* c, svpfo_state* st, double* h, double* h_prev, double* v,
               double y_t, double y_prev, const svpfo_noise* nz, svpfo_out* out);

/* T ticks with pre-drawn noise laid out as z_pred[T][n], xi[T][xi_rows][n],
 * u_pred/u_rej/z_rej[T][n] (NULL when unused).  Outputs are length T arrays
 * (any may be NULL). */
int svpfo_run(const svpfo_cfg* c, svpfo_state* st, double* h, double* h_prev, double* v,
              int T, const float* y, float y_prev0,
              const float* z_pred, const float* u_pred, const float* xi, int xi_rows,
              const float* u_rej, const float* z_rej,
              double* loglik, double* pred_loglik, double* vol, double* h_mean,
              double* ksd, double* bw, int* steps);

/* ---- the single-block fused step (src/svpf_fully_fused_v2.cu:90-712) ----
 * Scalar arguments of svpf_fully_fused_step_v2 (include/svpf_fully_fused_v2.cuh:43-116)
 * in the reference's order, as doubles and ints. */
typedef struct svpfo_v2_args {
    int n, timestep;
    int use_guide, use_guided_predict, use_guide_preserving;
    int use_newton, use_full_newton;
    int use_rejuvenation, use_adaptive_mu, use_adaptive_sigma;
    int n_stein_steps, n_anneal_steps, stein_sign_mode;
    int pad_;
    double y_t, y_prev, h_mean_prev, vol_prev, ksd_prev;
    double rho, sigma_z, mu, nu, lik_offset, student_t_const, student_t_implied_offset, gamma;
    double delta_rho, delta_sigma, mim_jump_prob, mim_jump_scale;
    double guide_strength_base, guide_strength_max, guide_innovation_threshold;
    double guided_alpha_base, guided_alpha_shock, guided_innovation_thresh_predict;
    double rejuv_ksd_threshold, rejuv_prob, rejuv_blend;
    double mu_ema_alpha, sigma_boost_threshold, sigma_boost_max;
    double step_size, temperature, rmsprop_rho, rmsprop_eps;
} svpfo_v2_args;

/* The call's device-pointer scalars: guide_mean / guide_var / mu_state are read
 * and written, the rest are outputs. */
typedef struct svpfo_v2_io {
    double guide_mean, guide_var, mu_state;
    double h_mean, vol, loglik, bandwidth, ksd;
} svpfo_v2_io;

/* One fused step.  h, h_prev, grad, log_w, v, inv_hess: length-n arrays (the
 * reference's h, h_prev, grad_log_p, log_weights, d_grad_v, d_inv_hessian).
 * Noise: with z_pred NULL every draw comes from the counter generator (predict
 * normal and MIM uniform: purpose 0; Langevin normal of iteration k: purpose
 * 16 + k; rejuvenation uniform and normal: purpose 1); otherwise z_pred[n],
 * u_pred[n] (NULL: no jump), xi[xi_rows][n] (row k feeds iteration k; NULL or
 * beyond xi_rows: 0), u_rej[n] (NULL: no rejuvenation), z_rej[n] (NULL: 0). */
int svpfo_v2_step(const svpfo_v2_args* a, double* h, double* h_prev, double* grad, double* log_w, double* v,
                  double* inv_hess, const svpfo_noise* nz, svpfo_v2_io* io);

/* ---- svpf_initialize's particle draw (src/svpf_optimized_graph.cu:73-97, 349-451) ----
 * h_i = clamp(mu + sd z_i, -15, 5), z_i the cosine normal of the counter generator
 * at (particle i, tick 0, purpose 255, filter 0) keyed by seed; sd the stationary
 * standard deviation sqrt(sigma^2 / (1 - rho^2 + 1e-6)), scaled by
 * sqrt(nu_s / (nu_s - 2)) (nu_s clamped to >= 2.5) for the Student-t state. */
void svpfo_init_particles(int n, double mu, double rho, double sigma_z, int student_t_state, double nu_state,
                          uint64_t seed, double* h);

/* ---- RTS-style smoothing of the reported h_mean (src/svpf_optimized_graph.cu:619-717) ---- */
typedef struct svpfo_smooth {
    int lag, output_lag, head, pad_;
    double h_mean[8], h_var[8], y[8];
} svpfo_smooth;
/* Stores (h_mean, h_var, y) of the tick whose pre-increment timestep is `timestep`,
 * then corrects the window backwards once it is full; returns the output h_mean. */
double svpfo_smooth_step(svpfo_smooth* s, int timestep, double h_mean, double h_var, double y_t, double mu,
                         double rho, double sigma_z);

/* ---- the coordinate-wise two-factor step (src/svpf_optimized_graph.cu:744-897,
 * src/svpf_two_factor.cu:29-468, include/svpf_two_factor.cuh:245-300) ---- */
typedef struct svpfo_2f_cfg {
    int n, n_stein_steps, n_anneal_steps, use_annealing;
    int use_svld, use_mim, use_guide, use_guide_preserving;
    int use_exact_gradient, stein_sign, pad0_, pad1_;
    double mu, nu, rho_fast, sigma_fast, rho_slow, sigma_slow;
    double bw_floor_fast, bw_floor_slow, guide_strength;
    double mim_jump_prob, mim_jump_scale, temperature, rmsprop_rho, rmsprop_eps;
} svpfo_2f_cfg;
typedef struct svpfo_2f_state {
    int timestep, guide_2f_initialized;
    double guide_mean_fast, guide_mean_slow, guide_var_fast, guide_var_slow;
    double mean_fast_prev, mean_slow_prev;   /* factor means after the previous tick (0 before the first) */
    double vol_prev, h_mean_prev;
} svpfo_2f_state;
typedef struct svpfo_2f_out {
    double loglik, pred_loglik, vol, h_mean, mean_fast, mean_slow, bw_fast, bw_slow;
    int steps, pad_;
} svpfo_2f_out;
/* Per-filter EKF of the two factors with the innovation split (include/svpf_two_factor.cuh:245-300). */
void svpfo_ekf_2f_update(const svpfo_2f_cfg* c, svpfo_2f_state* s, double y_t);
/* One tick.  hf, hs: factor particles; hfp, hsp: their previous values; vf, vs:
 * RMSProp states; all length n, updated in place.  Noise from the counter
 * generator keyed by (seed, filter_id, timestep): the fast predict normal and
 * the MIM uniform from purpose 0, the slow predict normal from purpose 2, the
 * Langevin normals of iteration k from purposes 16 + k (fast) and 1024 + k (slow). */
int svpfo_tick_2f(const svpfo_2f_cfg* c, svpfo_2f_state* st, double* hf, double* hs, double* hfp, double* hsp,
                  double* vf, double* vs, double y_t, uint64_t seed, uint32_t filter_id, svpfo_2f_out* out);
/* svpf_initialize with use_two_factor: factor particles from the stationary laws,
 * fast and slow from the two normals of one purpose-255 draw. */
void svpfo_init_2f(int n, double rho_fast, double sigma_fast, double rho_slow, double sigma_slow, uint64_t seed,
                   double* hf, double* hs);

/* ---- the joint state-parameter filter (svpf_joint/src/svpf_joint.cu,
 * svpf_joint/src/svpf_joint_kernels.cu): particles (h, mu~, rho~, sigma~) with
 * mu = mu~, rho = sigmoid(rho~), sigma = exp(sigma~), moved by a 4-D IMQ Stein
 * transport.  Configuration fields in the order of SVPFJointConfig
 * (svpf_joint/include/svpf_joint.cuh). ---- */
typedef struct svpfo_joint_cfg {
    int n, n_stein_steps;
    double step_h, step_mu, step_rho, step_sigma;
    double diffusion_mu, diffusion_rho, diffusion_sigma;
    double collapse_thresh_mu, collapse_thresh_rho, collapse_thresh_sigma;
    double mu_prior_mean, mu_prior_var, rho_prior_mean, rho_prior_var, sigma_prior_mean, sigma_prior_var;
    double nu, student_t_const, student_t_implied_offset, lik_offset, prior_weight;
} svpfo_joint_cfg;
/* Per-step outputs: SVPFJointDiagnostics plus the bandwidths. */
typedef struct svpfo_joint_out {
    double h_mean, vol_mean, mu_mean, rho_mean, sigma_mean, mu_std, rho_std, sigma_std;
    double std_mu_tilde, std_rho_tilde, std_sigma_tilde, ess, log_likelihood;
    double bw[4];
    int mu_collapsed, rho_collapsed, sigma_collapsed, pad_;
} svpfo_joint_out;
/* x: [4][n] = h | mu~ | rho~ | sigma~, hp: [n] previous h.  Initial draw from the
 * counter generator (purposes 253: h and mu~, 252: rho~ and sigma~; tick 0). */
void svpfo_joint_init(const svpfo_joint_cfg* c, uint64_t seed, double* x, double* hp);
/* One step at tick `tick`: predict normals from purposes 3 (mu~, rho~) and 4
 * (sigma~, h) of the counter generator. */
int svpfo_joint_step(const svpfo_joint_cfg* c, double* x, double* hp, double y_t, uint64_t seed, uint32_t tick,
                     svpfo_joint_out* out);
/* The per-particle score (svpf_joint_kernels.cu:134-320): g[4], log weight. */
void svpfo_joint_grad(const svpfo_joint_cfg* c, const double xj[4], double h_prev, double y_t, double g[4],
                      double* log_w);

/* Counter-based generator: Philox4x32-10 (Salmon et al., SC'11). */
void svpfo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Student-t state noise from the counter generator (particle i, tick, filter). */
double svpfo_student_t_draw(uint64_t seed, uint32_t filter, uint32_t tick, uint32_t i, double nu_state);
double svpfo_u01_from_bits(uint32_t x);
double svpfo_normal_from_bits(uint32_t a, uint32_t b);

#ifdef __cplusplus
}
#endif
#endif
