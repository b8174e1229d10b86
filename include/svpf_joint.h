/*
 * svpf_joint.h -- the joint state-parameter SVPF on gfx950.
 *
 * Each particle carries (h, mu~, rho~, sigma~) with mu = mu~, rho = sigmoid(rho~),
 * sigma = exp(sigma~), and one Stein transport moves state and parameters together
 * under a 4-D inverse-multiquadric kernel with one bandwidth per coordinate.
 *
 * Drop-in for svpf_joint/include/svpf_joint.cuh: the same configuration,
 * diagnostics and state structures (field names and order), the same functions
 * and argument meaning.  Differences a caller can observe:
 *  - one kernel launch runs a whole step (predict, bandwidths, every Stein
 *    iteration, diagnostics); the host waits only when `diag` is requested;
 *  - noise comes from a counter-based generator keyed by (seed, particle, tick,
 *    purpose), so d_rng is NULL;
 *  - every particle reads the positions of the previous Stein iteration;
 *  - the bandwidth variances are two-pass (mean, then centred squares);
 *  - up to SVPF_JOINT_MAX_PARTICLES particles (the reference's shared-memory
 *    staging stops near 2048 on its target);
 *  - svpf_joint_initialize computes the diagnostics of the initial particles, so
 *    svpf_joint_get_params / get_vol return defined values before the first step;
 *  - with n_stein_steps = 0 the log weights are those of the predicted particles.
 */
#ifndef SVPF_JOINT_H
#define SVPF_JOINT_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVPF_JOINT_DEFAULT_PARTICLES 512
#define SVPF_JOINT_DEFAULT_STEIN_STEPS 5
#define SVPF_JOINT_BLOCK_SIZE 256
#define SVPF_JOINT_MAX_PARTICLES 4096

/* svpf_joint/include/svpf_joint.cuh:42-84 */
typedef struct {
    int n_particles;
    int n_stein_steps;
    float step_h;
    float step_mu;
    float step_rho;
    float step_sigma;
    float diffusion_mu;
    float diffusion_rho;
    float diffusion_sigma;
    float collapse_thresh_mu;
    float collapse_thresh_rho;
    float collapse_thresh_sigma;
    float mu_prior_mean;
    float mu_prior_var;
    float rho_prior_mean;
    float rho_prior_var;
    float sigma_prior_mean;
    float sigma_prior_var;
    float nu;
    float student_t_const;
    float student_t_implied_offset;
    float lik_offset;
    float prior_weight;
} SVPFJointConfig;

/* svpf_joint/include/svpf_joint.cuh:89-119 */
typedef struct {
    float h_mean;
    float vol_mean;
    float mu_mean;
    float rho_mean;
    float sigma_mean;
    float mu_std;
    float rho_std;
    float sigma_std;
    float std_mu_tilde;
    float std_rho_tilde;
    float std_sigma_tilde;
    int mu_collapsed;
    int rho_collapsed;
    int sigma_collapsed;
    float ess;
    float log_likelihood;
} SVPFJointDiagnostics;

/* svpf_joint/include/svpf_joint.cuh:124-168.  The device arrays are views of one
 * allocation; d_rng is NULL (counter-based noise). */
typedef struct {
    float* d_h;
    float* d_h_prev;
    float* d_mu_tilde;
    float* d_rho_tilde;
    float* d_sigma_tilde;
    float* d_grad_h;
    float* d_grad_mu;
    float* d_grad_rho;
    float* d_grad_sigma;
    float* d_log_w;
    float bw_h;
    float bw_mu;
    float bw_rho;
    float bw_sigma;
    void* d_rng;
    float* d_temp;
    float* d_reduce_out;
    float* d_param_mean;
    float* d_param_std;
    float* d_std_unconstrained;
    int* d_collapse_flags;
    SVPFJointConfig config;
    int timestep;
    float y_prev;
    hipStream_t stream;
    /* not in the reference: */
    unsigned long long seed;
    int last_error; /* hipError_t of the last failed launch or copy, 0 if none */
    void* impl;
} SVPFJointState;

/* svpf_joint/src/svpf_joint.cu:21-72 */
SVPFJointConfig svpf_joint_default_config(void);
/* svpf_joint/src/svpf_joint.cu:78-130.  NULL when allocation fails or when
 * n_particles lies outside [1, SVPF_JOINT_MAX_PARTICLES]. */
SVPFJointState* svpf_joint_create(const SVPFJointConfig* config, hipStream_t stream);
/* svpf_joint/src/svpf_joint.cu:136-171 */
void svpf_joint_initialize(SVPFJointState* state, unsigned long long seed);
/* svpf_joint/src/svpf_joint.cu:196-376.  Prints the reference's collapse warning
 * when diag is given and a parameter's diversity has collapsed. */
void svpf_joint_step(SVPFJointState* state, float y_t, SVPFJointDiagnostics* diag);
/* svpf_joint/src/svpf_joint.cu:382-387 */
float svpf_joint_get_vol(const SVPFJointState* state);
/* svpf_joint/src/svpf_joint.cu:389-397 */
void svpf_joint_get_params(const SVPFJointState* state, float* mu, float* rho, float* sigma);
/* svpf_joint/src/svpf_joint.cu:403-425 */
void svpf_joint_destroy(SVPFJointState* state);

/* ---- not in the reference ---- */
/* Words per tick written by svpf_joint_run_sequence: the 16 words of
 * SVPFJointDiagnostics followed by bw_h, bw_mu, bw_rho, bw_sigma. */
#define SVPF_JOINT_RECW 20
/* T steps in one launch (no host round trip per step).  d_y: T observations in
 * device memory; d_records: NULL or T * SVPF_JOINT_RECW floats in device memory.
 * Equivalent to T calls of svpf_joint_step except that no collapse warning is
 * printed.  Returns a hipError_t value (0 on success). */
int svpf_joint_run_sequence_device(SVPFJointState* state, const float* d_y, int T, float* d_records);
/* Host copies of the particles: each pointer NULL or n floats. */
int svpf_joint_get_particles(const SVPFJointState* state, float* h, float* h_prev, float* mu_tilde,
                             float* rho_tilde, float* sigma_tilde);
/* Overwrite the particles (NULL leaves an array unchanged). */
int svpf_joint_set_particles(SVPFJointState* state, const float* h, const float* h_prev, const float* mu_tilde,
                             const float* rho_tilde, const float* sigma_tilde);
/* Dynamic LDS bytes of the step kernel at n particles. */
int svpf_joint_lds_bytes(int n);

#ifdef __cplusplus
}
#endif

#endif
