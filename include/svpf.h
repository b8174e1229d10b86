/*
 * svpf.h -- C ABI of the MI355X Stein variational particle filter.
 *
 * Drop-in for the reference's public step surface (include/svpf.cuh): same
 * function names, argument meaning and SVPFState field names, with
 * hipStream_t in place of cudaStream_t.  Each declaration cites the reference
 * declaration it replaces (paths relative to the reference repository root).
 *
 * Differences from the reference, all additive:
 *   - errors are reported: every call records a status readable through
 *     svpf_last_error(state); svpf_create returns NULL on any failure;
 *   - a feature combination a kernel path cannot serve is rejected with
 *     SVPF_RC_UNSUPPORTED instead of being silently ignored (the
 *     coordinate-wise two-factor step, use_two_factor, runs in one workgroup
 *     and so takes n_particles up to the single-workgroup capacity);
 *   - svpf_step_graph really replays a captured HIP graph per tick;
 *   - three configuration fields: bandwidth_mode, use_linear_beta,
 *     use_ksd_early_stop (defaults reproduce the reference).
 */
#ifndef SVPF_AMD_H
#define SVPF_AMD_H

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/svpf.cuh:47-60 */
#define SVPF_DEFAULT_PARTICLES 512
#define SVPF_DEFAULT_STEIN_STEPS 5
#define SVPF_DEFAULT_NU 5.0f
#define SVPF_STEIN_STEP_SIZE 0.1f
#define SVPF_H_MIN -15.0f
#define SVPF_H_MAX 5.0f
#define SVPF_SMOOTH_MAX_LAG 8

/* Largest n_particles of one filter.  Up to the single-workgroup capacity
 * (svpf_tile_capacity_query) a filter runs in one workgroup's LDS; larger
 * filters run on the grid path, a per-tick sequence of grid-wide kernels. */
#define SVPF_GRID_MAX_PARTICLES 16384

/* include/svpf.cuh:84-90 */
#define SVPF_STEIN_SIGN_LEGACY 0
#define SVPF_STEIN_SIGN_PAPER 1
#ifndef SVPF_STEIN_SIGN_DEFAULT
#define SVPF_STEIN_SIGN_DEFAULT SVPF_STEIN_SIGN_LEGACY
#endif

/* Bandwidth rules (bandwidth_mode). */
#define SVPF_BANDWIDTH_REFERENCE_EMA 0 /* src/svpf_opt_kernels.cu:1159-1235 */
#define SVPF_BANDWIDTH_MEDIAN 1        /* ell^2 = lower median of pairwise r^2 */
#define SVPF_BANDWIDTH_SILVERMAN 2     /* src/svpf_fully_fused_v2.cu:397-424 */

/* Return codes (svpf_last_error).  Prefixed SVPF_RC_ so that the SVPFStatus
 * names of svpf_diagnostics.h (SVPF_OK, SVPF_ERR_*) can be included beside them. */
#define SVPF_RC_OK 0
#define SVPF_RC_INVALID_ARGUMENT 1
#define SVPF_RC_UNSUPPORTED 2
#define SVPF_RC_HIP 3
#define SVPF_RC_NOT_INITIALIZED 4
#define SVPF_RC_OUT_OF_MEMORY 5
#define SVPF_RC_TIMEOUT 6 /* a cooperative kernel's cross-workgroup exchange timed out */
#define SVPF_RC_TOO_LARGE 7 /* above the fixed scratch of svpf_run_sequence_device_noalloc */

/* include/svpf.cuh:257-262 */
typedef struct {
    float rho;
    float sigma_z;
    float mu;
    float gamma;
} SVPFParams;

struct svpf_impl;

/* include/svpf.cuh:290-547.  Configuration is by direct field writes between
 * calls, as in the reference; every field below is read on each tick. */
typedef struct {
    /* device arrays [n_particles] (include/svpf.cuh:292-305) */
    float* h;
    float* h_prev;
    float* h_pred;
    float* grad_log_p;
    float* log_weights;
    float* d_grad_v;

    int n_particles;
    int n_stein_steps;
    float nu;
    float student_t_const;
    float student_t_implied_offset;
    int timestep;
    float y_prev;
    hipStream_t stream;

    float lik_offset;
    int use_exact_gradient;

    int use_svld;
    int use_annealing;
    int use_adaptive_beta;
    int n_anneal_steps;
    float temperature;
    float rmsprop_rho;
    float rmsprop_eps;

    int use_mim;
    float mim_jump_prob;
    float mim_jump_scale;

    int use_local_params;
    float delta_rho;
    float delta_sigma;

    int use_asymmetric_rho;
    float rho_up;
    float rho_down;

    int use_newton;
    int use_full_newton;

    int use_guided;
    float guided_alpha_base;
    float guided_alpha_shock;
    float guided_innovation_threshold;

    int use_rejuvenation;
    float rejuv_ksd_threshold;
    float rejuv_prob;
    float rejuv_blend;

    int use_guide;
    int use_guide_preserving;
    float guide_strength;
    float guide_mean;
    float guide_var;
    float guide_K;
    int guide_initialized;

    int use_adaptive_guide;
    float guide_strength_base;
    float guide_strength_max;
    float guide_innovation_threshold;
    float vol_prev;

    int use_adaptive_mu;
    float mu_state;
    float mu_var;
    float mu_process_var;
    float mu_obs_var_scale;
    float mu_min;
    float mu_max;

    int use_adaptive_sigma;
    float sigma_boost_threshold;
    float sigma_boost_max;
    float sigma_z_effective;

    int stein_repulsive_sign;
    int use_fan_mode;

    int stein_min_steps;
    int stein_max_steps;
    float ksd_improvement_threshold;
    float ksd_prev;
    int stein_steps_used;

    int use_student_t_state; /* single-workgroup kernel (n < 256) or grid path; not the cooperative kernel */
    float nu_state;

    int use_smoothing;
    int smooth_lag;
    int smooth_output_lag;
    int smooth_head;
    float smooth_h_mean[SVPF_SMOOTH_MAX_LAG];
    float smooth_h_var[SVPF_SMOOTH_MAX_LAG];
    float smooth_y[SVPF_SMOOTH_MAX_LAG];

    int use_persistent_kernel; /* 1: cooperative multi-workgroup kernel (default for n >= 256) */
    int persistent_kernel_supported;
    int use_heun; /* single-workgroup kernel (n < 256) or grid path; not the cooperative kernel */

    /* include/svpf.cuh:533-544: the coordinate-wise two-factor step
     * (src/svpf_optimized_graph.cu:744-897), h = mu + h_fast + h_slow.  One
     * workgroup, n_particles up to the single-workgroup capacity
     * (svpf_tile_capacity_query); above it the step reports
     * SVPF_RC_UNSUPPORTED.  guide_* are read and written by each tick. */
    int use_two_factor;
    float rho_fast;
    float sigma_fast;
    float rho_slow;
    float sigma_slow;
    float bw_floor_fast;
    float bw_floor_slow;
    float guide_mean_fast;
    float guide_mean_slow;
    float guide_var_fast;
    float guide_var_slow;
    int guide_2f_initialized;

    /* extensions (defaults reproduce the reference) */
    int bandwidth_mode;
    int use_linear_beta;
    int use_ksd_early_stop;

    struct svpf_impl* impl;
} SVPFState;

/* include/svpf.cuh:552-558 */
typedef struct {
    float log_lik_increment;
    float vol_mean;
    float vol_std;
    float h_mean;
    float mu_estimate;
} SVPFResult;

/* include/svpf.cuh:573 -- NULL on any failure (reference: malloc failure only),
 * including n_particles above SVPF_GRID_MAX_PARTICLES. */
SVPFState* svpf_create(int n_particles, int n_stein_steps, float nu, hipStream_t stream);
/* include/svpf.cuh:578 */
void svpf_destroy(SVPFState* state);
/* include/svpf.cuh:587 -- particles drawn by the counter-based generator keyed by seed. */
void svpf_initialize(SVPFState* state, const SVPFParams* params, unsigned long long seed);
/* include/svpf.cuh:597 -- uses state->y_prev as y_{t-1}. */
void svpf_step(SVPFState* state, float y_t, const SVPFParams* params, SVPFResult* result);
/* include/svpf.cuh:602-603 -- rng_seed re-keys the generator for this tick. */
void svpf_step_seeded(SVPFState* state, float y_t, const SVPFParams* params, unsigned long long rng_seed,
                      SVPFResult* result);
/* include/svpf.cuh:684-691 */
void svpf_run_sequence(SVPFState* state, const float* h_observations, int T, const SVPFParams* params,
                       float* h_loglik_out, float* h_vol_out);
/* include/svpf.cuh:701-708 -- a device loop: one launch advances all T ticks. */
void svpf_run_sequence_device(SVPFState* state, const float* d_observations, int T, const SVPFParams* params,
                              float* d_loglik_out, float* d_vol_out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory during a call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): svpf_run_sequence_device with the per-tick
 * records in a fixed scratch array compiled into the library (one per GPU, held
 * until the process ends) instead of a per-state buffer grown with hipMalloc.
 * The call allocates nothing; T above SVPF_NOALLOC_MAX_TICKS is rejected with
 * SVPF_RC_TOO_LARGE.  Calls that overlap on one GPU share that array and race:
 * issue them from one host thread, one at a time per GPU. */
#define SVPF_NOALLOC_MAX_TICKS 32768
void svpf_run_sequence_device_noalloc(SVPFState* state, const float* d_observations, int T,
                                      const SVPFParams* params, float* d_loglik_out, float* d_vol_out);
/* include/svpf.cuh:767-775 */
void svpf_step_adaptive(SVPFState* state, float y_t, float y_prev, const SVPFParams* params, float* h_loglik_out,
                        float* h_vol_out, float* h_mean_out);
/* include/svpf.cuh:843-851 */
void svpf_step_graph(SVPFState* state, float y_t, float y_prev, const SVPFParams* params, float* h_loglik_out,
                     float* h_vol_out, float* h_mean_out);
/* include/svpf.cuh:856 */
bool svpf_graph_is_captured(SVPFState* state);
/* include/svpf.cuh:869 */
void svpf_graph_invalidate(SVPFState* state);
/* include/svpf.cuh:795 */
void svpf_get_particles(const SVPFState* state, float* h_out);
/* include/svpf.cuh:800 */
void svpf_get_stats(const SVPFState* state, float* h_mean, float* h_std);
/* include/svpf.cuh:811 */
float svpf_get_ess(const SVPFState* state);
/* src/svpf_optimized_graph.cu:1490-1493 (defined there, never declared) */
void svpf_get_ksd_stats(const SVPFState* state, float* ksd_out, int* steps_used_out);

/* No reference counterpart: status of the last call on this state (SVPF_RC_OK ...)
 * and a readable description (the HIP error string for SVPF_RC_HIP). */
int svpf_last_error(const SVPFState* state);
const char* svpf_last_error_string(const SVPFState* state);

/* No reference counterpart: parity hooks for identical-input comparisons.
 * svpf_set_particles uploads h, h_prev and the RMSProp second moment (any may
 * be NULL: h_prev defaults to h, v to zero).  svpf_set_predrawn_noise points
 * the step at device arrays of pre-drawn noise laid out [T][n]
 * (z_pred, u_pred, u_rej, z_rej) and [T][xi_rows][n] (xi); tick k of the
 * filter reads row k - t_base.  Passing z_pred = NULL returns to the
 * counter-based generator.  svpf_set_scalar_state sets the device scalars
 * the reference keeps on the device (mean of h after the last tick and the
 * EMA bandwidth state). */
int svpf_set_particles(SVPFState* state, const float* h, const float* h_prev, const float* v);
int svpf_get_rmsprop_state(const SVPFState* state, float* v_out);
int svpf_set_predrawn_noise(SVPFState* state, const float* d_z_pred, const float* d_u_pred, const float* d_xi,
                            int xi_rows, const float* d_u_rej, const float* d_z_rej, int t_base);
int svpf_set_scalar_state(SVPFState* state, float h_mean_prev, float bw_sq_ema, float return_ema,
                          float return_var);
/* No reference counterpart: per-tick diagnostics of the last tick
 * (loglik, vol, h_mean, ksd, bandwidth, predictive loglik, steps, ell^2). */
int svpf_get_last_outputs(const SVPFState* state, float out8[8]);
/* The factor particles of a two-factor filter (include/svpf.cuh:235-246 keeps
 * them as the device arrays d_h_fast / d_h_slow of the optimized backend):
 * copied to host arrays of n_particles floats, either may be NULL.  The step
 * also writes h = mu + h_fast + h_slow into state->h, which svpf_get_particles
 * reads. */
int svpf_get_two_factor_particles(const SVPFState* state, float* h_fast, float* h_slow);

/* No reference counterpart: shape of the cooperative multi-workgroup kernel
 * that serves one filter when use_persistent_kernel is set (the default for
 * n_particles >= 256).  lanes_per_row in {1,2,4,8,16,32} gives 256 /
 * lanes_per_row particles per workgroup; xcd_stride 8 places all workgroups
 * on one XCD, 1 spreads them.  Pass -1 for the defaults. */
int svpf_set_coop_shape(SVPFState* state, int lanes_per_row, int xcd_stride);
int svpf_get_coop_shape(const SVPFState* state, int* lanes_per_row, int* workgroups, int* xcd_stride);
/* No reference counterpart: the cooperative kernel runs as a plain launch
 * (its workgroups are co-resident on an otherwise free GPU); when an exchange
 * times out -- nothing is written back then -- the step is repeated with a
 * cooperative launch (co-residency guaranteed), which the filter keeps.
 * svpf_coop_launch_is_cooperative reports 1 once that happened.
 * svpf_set_coop_spin_limit bounds each exchange wait at `polls` polls (0: the
 * default 2^20, at most 2^24); a negative value makes every exchange wait for
 * a tag nobody publishes and time out after -polls polls (failure-path
 * testing, at least -2^20). */
int svpf_coop_launch_is_cooperative(const SVPFState* state);
int svpf_set_coop_spin_limit(SVPFState* state, int polls);
/* No reference counterpart: with the median bandwidth (bandwidth_mode
 * SVPF_BANDWIDTH_MEDIAN) the cooperative kernel runs one more workgroup that
 * sorts the predicted particles and selects the median pairwise distance while
 * the row workgroups compute the first Stein iteration's gradient and exchange
 * its records; on = 0 makes every row workgroup select the median itself
 * instead.  Results are identical either way.  Default on. */
int svpf_set_coop_median_workgroup(SVPFState* state, int on);
/* Failure-path testing: the given workgroup of the cooperative kernel (-1: none)
 * gives up the waits of a launch's final exchange after 4 polls, after publishing
 * its rows, while the others complete; the launch then commits nothing. */
int svpf_set_coop_fault(SVPFState* state, int workgroup);

/* No reference counterpart: the kernel path serving one filter.
 *   SVPF_PATH_SINGLE  one workgroup holds the filter in LDS;
 *   SVPF_PATH_COOP    the cooperative multi-workgroup kernel (use_persistent_kernel);
 *   SVPF_PATH_GRID    a per-tick sequence of grid-wide kernels (2 <= n_particles
 *                     <= SVPF_GRID_MAX_PARTICLES; the only path above the
 *                     single-workgroup capacity).
 * svpf_set_grid_path: SVPF_GRID_AUTO (default) picks by n_particles and the
 * features in use -- cooperative kernel for 256 <= n < 1024 where it applies,
 * grid path from n = 1024 up and for 256 <= n < 1024 otherwise, one workgroup
 * below 256; SVPF_GRID_ON forces the grid path; SVPF_GRID_OFF forbids it
 * (SVPF_RC_UNSUPPORTED when n_particles exceeds the single-workgroup
 * capacity).  svpf_get_path returns the SVPF_PATH_* of the next step, -1 for a
 * bad state. */
#define SVPF_PATH_SINGLE 0
#define SVPF_PATH_COOP 1
#define SVPF_PATH_GRID 2
#define SVPF_GRID_AUTO (-1)
#define SVPF_GRID_OFF 0
#define SVPF_GRID_ON 1
int svpf_set_grid_path(SVPFState* state, int mode);
int svpf_get_path(const SVPFState* state);

/* include/svpf.cuh:626-632 */
static inline void svpf_set_stein_sign_mode(SVPFState* state, int mode) {
    state->stein_repulsive_sign = (mode == SVPF_STEIN_SIGN_PAPER) ? 1 : 0;
}
/* include/svpf.cuh:638-640 */
static inline int svpf_get_stein_sign_mode(const SVPFState* state) { return state->stein_repulsive_sign; }
/* include/svpf.cuh:661-667 */
static inline void svpf_set_fan_mode(SVPFState* state, int enable) {
    state->use_fan_mode = enable ? 1 : 0;
    if (enable) state->stein_repulsive_sign = SVPF_STEIN_SIGN_PAPER;
}
/* include/svpf.cuh:673-675 */
static inline int svpf_get_fan_mode(const SVPFState* state) { return state->use_fan_mode; }

/* ------------------------------------------------------------------------ */
/* Batched form (no reference counterpart: the reference runs one filter per
 * SVPFState).  B independent filters with their own particle counts and
 * parameters advance through ticks [t0, t1) in persistent launches, one per
 * tile group.  All pointers are device pointers except group arrays.  The
 * record layouts are those of newton-stein-variational-particle-filter_amd/
 * csrc/svpf_types.h (svpf_fparams, svpf_fstate, svpf_fdesc).  order lists
 * filter indices grouped by tile; group g covers order[group_start[g] ..
 * group_start[g+1]) and runs with tile group_tile[g] and block size
 * group_bs[g] on grid group_grid[g] workgroups.  queues holds one int per
 * group, then one per filter (n_groups + group_start[n_groups] ints; the call
 * resets them): a group's work items are the ticks of its filters, interleaved,
 * and a filter's word counts its finished ticks (a tick whose wait for the
 * previous one exceeds about a second of polling sets bit 1 of the filter's
 * status).  A group holding a filter with the Heun step or the Student-t state
 * (SVPF_F_HEUN / SVPF_F_STUDENT_STATE) ORs SVPF_TILE_EXTENDED into its tile id;
 * without it such a filter is rejected (its svpf_fstate status is set). */
int svpf_batched_run(const void* d_params, void* d_fstate, const void* d_desc, float* d_pool_h, float* d_pool_hp,
                     float* d_pool_v, const int* d_order, const int* group_start, const int* group_tile,
                     const int* group_bs, const int* group_grid, const int* group_nmax, int n_groups, int* d_queues,
                     int t0, int t1, hipStream_t stream);
/* ------------------------------------------------------------------------ */
/* Differentiable form (no reference counterpart: the reference has no
 * adjoint; its PyTorch path differentiates by autograd,
 * py-torch/diff_svpf_torch.py:147-217).  B filters of n particles (2..2048)
 * with the plain feature set and per-filter theta = (mu, rho, sigma_z, nu).
 * config is a host pointer to one svpf_dconfig, d_desc a device array of B
 * svpf_ddesc (csrc/svpf_types.h).  svpf_diff_forward writes the predictive
 * log-likelihood of every tick, the median bandwidth of every tick and a
 * checkpoint (x, v) every checkpoint_every ticks; svpf_diff_backward takes
 * dL/d pred_ll and writes dL/d theta (and dL/d h0), recomputing each
 * checkpoint segment into the tape.  Both return SVPF_RC_OK or an SVPF_ERR_*
 * code.  svpf_diff_tape_floats is the tape size per filter.  With the optional
 * svpf_ddesc.drift ([T][S][n] floats per filter) the forward stores every
 * iteration's Stein drift and the backward replays positions from it instead
 * of recomputing the transport. */
size_t svpf_diff_tape_floats(int n_particles, int n_stein, int checkpoint_every);
int svpf_diff_forward(const void* config, const void* d_desc, int B, hipStream_t stream);
int svpf_diff_backward(const void* config, const void* d_desc, int B, hipStream_t stream);

/* Two-factor form (no reference counterpart as a 2-D filter: the reference's
 * two-factor branch, src/svpf_optimized_graph.cu:744-897, moves each factor
 * with its own 1-D kernel).  B filters of n particles (2..2048) with
 * x = (f, s), h = mu + f + s, a 2-D IMQ kernel and a 2x2 Newton
 * preconditioner; theta = (mu, rho_f, sigma_f, rho_s, sigma_s, nu).  Records:
 * svpf_d2config (host pointer) and svpf_d2desc (device array of B),
 * csrc/svpf_types.h.  The forward writes per-tick records (predictive
 * loglik, h / f / s means, vol, bandwidths) and checkpoints; the backward
 * takes dL/d pred_ll and writes dL/d theta and dL/d x0. */
size_t svpf_diff2d_tape_floats(int n_particles, int n_stein, int checkpoint_every);
int svpf_diff2d_forward(const void* config, const void* d_desc, int B, hipStream_t stream);
int svpf_diff2d_backward(const void* config, const void* d_desc, int B, hipStream_t stream);

/* Largest particle count a tile of the given block size holds (0: unknown tile).
 * Block sizes 256 and 512 take tiles 0-7, 1024 takes tiles 0-4, each with or
 * without SVPF_TILE_EXTENDED. */
#define SVPF_TILE_EXTENDED 0x100
int svpf_tile_capacity_query(int tile, int block_size);
/* Workgroups of that tile resident per CU for filters of up to nmax particles
 * (HIP occupancy of the kernel with its LDS; 0: unknown tile).  Sizes the
 * persistent grid of svpf_batched_run. */
int svpf_tile_occupancy_query(int tile, int block_size, int nmax);
/* Dynamic LDS bytes a workgroup needs for filters of up to nmax particles. */
size_t svpf_lds_bytes_query(int nmax);
/* Library version string. */
const char* svpf_version(void);
/* No counterpart in the reference: the bounded digamma the adjoint and score
 * kernels use (the same code, compiled for the host), for host-side checks.
 * NaN for a non-positive or non-finite argument. */
float svpf_digamma_f32(float x);

/* ---- Score diagnostics beside the step (include/svpf.cuh:895-1213) ----
 * Posterior-weighted scores of the SV model read from the filter after a
 * step: d/dnu of the Student-t observation density at the particles h, and
 * d/dsigma of the transition density of h_pred given h_prev, each with its
 * weighted mean z^2 or eps^2/sigma^2 (csrc/svpf_diag.hip, one reduction
 * kernel on the state's stream and a synchronous read-back).  The host side
 * keeps the reference's EMA smoothing, CSV log and shock state machine.  The
 * reference declares svpf_compute_all_gradients (include/svpf.cuh:1102) and
 * svpf_test_nu_gradient_synthetic without defining them in the library; they
 * are not part of this ABI either. */

/* include/svpf.cuh:895-929 */
typedef struct {
    float* d_nu_grad;   /* device: weighted mean nu score */
    float* d_z_sq_mean; /* device: weighted mean z^2 */
    float* d_mu_grad;   /* unused (null), as there */
    float* d_rho_grad;
    float* d_sigma_grad;
    float* d_fisher;
    float* d_fisher_inv;
    float nu_gradient_ema; /* EMA, alpha 0.1 */
    float z_sq_ema;
    float mu_gradient_ema;
    float rho_gradient_ema;
    float sigma_gradient_ema;
    int shock_state; /* 0 calm, 1 shock, 2 recovery */
    int ticks_in_state;
    float shock_threshold;         /* z^2 entering SHOCK (9) */
    int shock_duration;            /* ticks in SHOCK (20) */
    int recovery_duration;         /* ticks in RECOVERY before CALM (50) */
    float recovery_exit_threshold; /* z^2 below which RECOVERY may end (4) */
    bool enable_logging;
    FILE* log_file;
    bool initialized;
} SVPFGradientDiagnostics;

/* include/svpf.cuh:942-947: mu, rho = tanh(eta), sigma = exp(kappa), nu = 2 + exp(kappa_nu) */
typedef struct {
    float mu;
    float eta;
    float kappa;
    float kappa_nu;
} SVPFThetaUnconstrained;

/* include/svpf.cuh:958-981 */
typedef struct {
    SVPFThetaUnconstrained theta;
    float F[4][4];
    float F_ema_decay;
    float F_reg;
    float base_lr;
    float lr_shock_mult;
    float grad_clip;
    float prior_weight;
    SVPFThetaUnconstrained theta_prior;
    int warmup_ticks;
    bool learning_enabled;
} SVPFNaturalGradientTuner;

/* include/svpf.cuh:994; NULL on allocation failure */
SVPFGradientDiagnostics* svpf_gradient_diagnostic_create(bool enable_logging, const char* log_path);
/* include/svpf.cuh:999 */
void svpf_gradient_diagnostic_destroy(SVPFGradientDiagnostics* diag);
/* include/svpf.cuh:1019-1026: the nu score at y_t over the current particles,
 * weighted by the log-weights, into diag's device buffers and EMAs (and its log) */
void svpf_compute_nu_diagnostic(SVPFState* state, SVPFGradientDiagnostics* diag, float y_t, int timestep,
                                float* nu_grad_out, float* z_sq_mean_out);
/* include/svpf.cuh:1034-1039: the same without diagnostic state (the state's
 * own scratch, no allocation) */
void svpf_compute_nu_diagnostic_simple(SVPFState* state, float y_t, float* nu_grad_out, float* z_sq_mean_out);
/* include/svpf.cuh:1062-1067: the sigma score over h_pred given h_prev, at the
 * state's effective sigma (params->sigma_z when that is below 1e-6) */
void svpf_compute_sigma_diagnostic_simple(SVPFState* state, const SVPFParams* params, float* sigma_grad_out,
                                          float* eps_sq_norm_out);
/* include/svpf.cuh:1074-1077: device-to-device copy of h on the state's stream */
void svpf_snapshot_particles(SVPFState* state, float* d_h_buffer);
/* include/svpf.cuh:1125-1143 */
void svpf_update_shock_state(SVPFGradientDiagnostics* diag, float z_sq);
int svpf_get_shock_state(const SVPFGradientDiagnostics* diag);
bool svpf_should_learn(const SVPFGradientDiagnostics* diag);
float svpf_get_lr_multiplier(const SVPFGradientDiagnostics* diag, float lr_shock_mult);
/* include/svpf.cuh:1157-1198 (the update is the reference's placeholder: it
 * returns the current parameters) */
SVPFNaturalGradientTuner* svpf_tuner_create(const SVPFParams* params, float nu, float base_lr, float prior_weight);
void svpf_tuner_destroy(SVPFNaturalGradientTuner* tuner);
void svpf_tuner_update(SVPFNaturalGradientTuner* tuner, const SVPFGradientDiagnostics* diag, SVPFParams* params_out,
                       float* nu_out);
void svpf_tuner_get_params(const SVPFNaturalGradientTuner* tuner, SVPFParams* params_out, float* nu_out);

/* include/svpf.cuh:1219-1256 */
static inline float svpf_constrain_rho(float eta) { return tanhf(eta); }
static inline float svpf_unconstrain_rho(float rho) {
    rho = fminf(fmaxf(rho, -0.999f), 0.999f);
    return 0.5f * logf((1.0f + rho) / (1.0f - rho));
}
static inline float svpf_constrain_sigma(float kappa) { return expf(kappa); }
static inline float svpf_unconstrain_sigma(float sigma) { return logf(fmaxf(sigma, 1e-8f)); }
static inline float svpf_constrain_nu(float kappa_nu) { return 2.0f + expf(kappa_nu); }
static inline float svpf_unconstrain_nu(float nu) { return logf(fmaxf(nu - 2.0f, 1e-8f)); }
static inline float svpf_drho_deta(float rho) { return 1.0f - rho * rho; }
static inline float svpf_dsigma_dkappa(float sigma) { return sigma; }
static inline float svpf_dnu_dkappa_nu(float nu) { return nu - 2.0f; }

#ifdef __cplusplus
}
#endif
#endif
